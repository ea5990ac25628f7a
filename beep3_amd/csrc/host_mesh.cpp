// Mesh generation, node patches and quadrature point sets (host C++).
//
// Node patches follow BasicNodePatch (src/bem/node_patch.cpp:55-130): one
// patch per vertex made of the barycentric-dual thirds of the adjacent planar
// triangles; area = sum of 4-point galerkin weights, centroid = weighted mean
// of those points, normal = normalised weighted mean of triangle normals.
// Point sets follow BasicTriangle::get_constant_basis_quad_points
// (src/bem/triangle.h:388-479).
#include "host.hpp"
#include "curved.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <stdexcept>
#include <unordered_map>

namespace beep {

// Rule tables.  The values are the reference's (src/bem/triangle.h:75-96):
// Strang-Fix 6-point, 1-point centroid, 4-point (negative centroid weight),
// 7-point Radon rule as rounded there, and the 16-point degree-8 rule.
static Rule make_rule(int order) {
    Rule r;
    r.order = order;
    switch (order) {
        case 1:
            r.u = {1.0 / 3.0};
            r.v = {1.0 / 3.0};
            r.w = {1.0};
            break;
        case 4:
            r.u = {1.0 / 3.0, 0.2, 0.6, 0.2};
            r.v = {1.0 / 3.0, 0.6, 0.2, 0.2};
            r.w = {-0.5625, 0.520833333333, 0.520833333333, 0.520833333333};
            break;
        case 6: {
            const double a = 0.659027622374092, b = 0.231933368553031, c = 0.109039009072877;
            r.u = {a, a, b, b, c, c};
            r.v = {b, c, a, c, a, b};
            r.w.assign(6, 0.16666666666666666667);
            break;
        }
        case 7: {
            const double a1 = 0.05971587, b1 = 0.47014206, a2 = 0.79742698, b2 = 0.10128650;
            r.u = {1.0 / 3.0, a1, b1, b1, a2, b2, b2};
            r.v = {1.0 / 3.0, b1, a1, b1, b2, a2, b2};
            const double w1 = 0.132394153333333, w2 = 0.125939186666666;
            r.w = {0.225, w1, w1, w1, w2, w2, w2};
            break;
        }
        case 16: {
            const double c0 = 0.333333333333333;
            const double a1 = 0.081414823414554, b1 = 0.459292588292723;
            const double a2 = 0.65886138449648, b2 = 0.17056930775176;
            const double a3 = 0.898905543365938, b3 = 0.050547228317031;
            const double a4 = 0.008394777409958001, b4 = 0.263112829634638, c4 = 0.728492392955404;
            r.u = {c0, a1, b1, b1, a2, b2, b2, a3, b3, b3, a4, b4, c4, b4, c4, a4};
            r.v = {c0, b1, b1, a1, b2, b2, a2, b3, b3, a3, b4, c4, a4, a4, b4, c4};
            const double w0 = 0.144315607677787, w1 = 0.09509163426728499, w2 = 0.103217370534718,
                         w3 = 0.032458497623198, w4 = 0.027230314174435;
            r.w = {w0, w1, w1, w1, w2, w2, w2, w3, w3, w3, w4, w4, w4, w4, w4, w4};
            break;
        }
        default:
            throw std::runtime_error("unsupported quadrature rule order");
    }
    return r;
}

const Rule& rule_for_order(int order) {
    static const Rule r1 = make_rule(1), r4 = make_rule(4), r6 = make_rule(6), r7 = make_rule(7),
                      r16 = make_rule(16);
    switch (order) {
        case 1: return r1;
        case 4: return r4;
        case 6: return r6;
        case 7: return r7;
        case 16: return r16;
    }
    throw std::runtime_error("unsupported quadrature rule order");
}

Vec3 Quat::rotate(const Vec3& v) const {
    // v' = q v q*
    Vec3 u{x, y, z};
    Vec3 t = u.cross(v) * 2.0;
    return v + t * w + u.cross(t);
}

// ---- curved patches (PNG1) ----
namespace {

// pt projected onto the plane through `origin` with normal n, then pulled to a
// third of its original distance from origin along that direction
Vec3 tangent_third(const Vec3& pt, const Vec3& origin, const Vec3& n) {
    const Vec3 proj = pt + n * n.dot(origin - pt);
    return origin + (proj - origin).unit() * ((origin - pt).norm() / 3.0);
}

// In-plane frame of triangle (a, b, c): origin a, x along b - a (length
// `scale`), y = unit(planar normal x x); (dx, dy) are c's coordinates.
struct Frame {
    Vec3 o, x, y;
    double scale, dx, dy;
    Frame(const Vec3& a, const Vec3& b, const Vec3& c) {
        o = a;
        x = b - a;
        scale = x.norm();
        x = x / scale;
        const Vec3 n = (b - a).cross(c - a).unit();
        y = n.cross(x).unit();
        dx = (c - a).dot(x);
        dy = (c - a).dot(y);
    }
};

// Third vertex of each triangle carried into the other's frame, scaled so the
// two base edges match: (first's third vertex in the second, second's third
// vertex in the first)
void swap_thirds(const Frame& f1, const Frame& f2, Vec3& one_in_two, Vec3& two_in_one) {
    one_in_two = f2.o + (f2.x * f1.dx + f2.y * f1.dy) * (f2.scale / f1.scale);
    two_in_one = f1.o + (f1.x * f2.dx + f1.y * f2.dy) * (f1.scale / f2.scale);
}

// Third vertex of frame f placed on the base a -> b in the plane with normal n
Vec3 third_on(const Frame& f, const Vec3& a, const Vec3& b, const Vec3& n) {
    Vec3 x = b - a;
    const double scale = x.norm();
    x = x / scale;
    const Vec3 y = n.cross(x).unit();
    return a + (x * f.dx + y * f.dy) * (scale / f.scale);
}

}  // namespace

CurvedTri::CurvedTri(const Vec3& v1, const Vec3& v2, const Vec3& v3, const Vec3& n1, const Vec3& n2,
                     const Vec3& n3, const Vec3& n12, const Vec3& n23, const Vec3& n31) {
    const Vec3 &b300 = v1, &b030 = v2, &b003 = v3;
    const Vec3 p1_210 = tangent_third(v2, v1, n1);
    const Vec3 p1_201 = tangent_third(v3, v1, n1);
    const Vec3 p2_120 = tangent_third(v1, v2, n2);
    const Vec3 p2_021 = tangent_third(v3, v2, n2);
    const Vec3 p3_102 = tangent_third(v1, v3, n3);
    const Vec3 p3_012 = tangent_third(v2, v3, n3);
    Vec3 p2_201, p2_102, p3_120, p3_210, p1_012, p1_021;
    swap_thirds(Frame(b300, p1_210, p1_201), Frame(p2_120, b030, p2_021), p1_021, p2_201);
    swap_thirds(Frame(b030, p2_021, p2_120), Frame(p3_012, b003, p3_102), p2_102, p3_120);
    swap_thirds(Frame(b003, p3_102, p3_012), Frame(p1_201, b300, p1_210), p3_210, p1_012);
    const Vec3 nt = (v2 - v1).cross(v3 - v1).unit();
    const Vec3 a12 = (nt + n12).unit(), a23 = (nt + n23).unit(), a31 = (nt + n31).unit();
    const Vec3 c12_1 = third_on(Frame(b300, p1_210, p1_201), p1_210, p2_120, a12);
    const Vec3 c12_2 = third_on(Frame(p2_120, b030, p2_021), p1_210, p2_120, a12);
    const Vec3 c23_2 = third_on(Frame(b030, p2_021, p2_120), p2_021, p3_012, a23);
    const Vec3 c23_3 = third_on(Frame(p3_012, b003, p3_102), p2_021, p3_012, a23);
    const Vec3 c31_3 = third_on(Frame(b003, p3_102, p3_012), p3_102, p1_201, a31);
    const Vec3 c31_1 = third_on(Frame(p1_201, b300, p1_210), p3_102, p1_201, a31);
    const Vec3* order[21] = {&b300,   &b030,   &b003,   &p1_210, &p1_201, &p2_120, &p2_021,
                             &p3_102, &p3_012, &p2_201, &p2_102, &p3_120, &p3_210, &p1_012,
                             &p1_021, &c12_1,  &c12_2,  &c23_2,  &c23_3,  &c31_3,  &c31_1};
    for (int i = 0; i < 21; ++i) {
        ctl[3 * i] = order[i]->x;
        ctl[3 * i + 1] = order[i]->y;
        ctl[3 * i + 2] = order[i]->z;
    }
}

Vec3 CurvedTri::point(double u, double v) const {
    double p[3], du[3], dv[3];
    beepcurve::eval(ctl, u, v, p, du, dv);
    return Vec3(p[0], p[1], p[2]);
}

void CurvedTri::tangents(double u, double v, Vec3& du, Vec3& dv) const {
    double p[3], a[3], b[3];
    beepcurve::eval(ctl, u, v, p, a, b);
    du = Vec3(a[0], a[1], a[2]);
    dv = Vec3(b[0], b[1], b[2]);
}

// Vertex normals (Vertex::set_normal, src/bem/vertex.cpp:86-141: the
// area-weighted mean of the planar normals around the vertex, triangles below
// 1e-12 in area left out) and the planar normals across each edge
// (Triangle::set_adjacent_normal, src/bem/triangle.h:737-757, over the
// mesh's edges, src/bem/mesh.cpp:997-1004); then one curved triangle each.
static void build_curved(Mesh& m) {
    const size_t nt = m.tris.size();
    std::vector<Vec3> tn(nt);
    std::vector<double> ta(nt);
    for (size_t t = 0; t < nt; ++t) {
        const auto& tr = m.tris[t];
        const Vec3 cr = (m.verts[tr[1]] - m.verts[tr[0]]).cross(m.verts[tr[2]] - m.verts[tr[0]]);
        ta[t] = 0.5 * cr.norm();
        tn[t] = cr.unit();
    }
    std::vector<Vec3> vn(m.verts.size());
    for (size_t v = 0; v < m.verts.size(); ++v) {
        Vec3 s;
        for (int k = m.vt_ptr[v]; k < m.vt_ptr[v + 1]; ++k) {
            const int t = m.vt_tri[k];
            if (ta[t] >= 1e-12) s += tn[t] * ta[t];
        }
        if (s.dot(s) < 1e-12) throw std::runtime_error("curved patches: a vertex has no usable triangle");
        vn[v] = s.unit();
    }
    // edge (lo, hi) -> the triangles on it
    std::unordered_map<uint64_t, std::array<int, 2>> edge;
    edge.reserve(nt * 2);
    auto ekey = [](int a, int b) {
        return (uint64_t(std::min(a, b)) << 32) | uint64_t(uint32_t(std::max(a, b)));
    };
    for (size_t t = 0; t < nt; ++t)
        for (int k = 0; k < 3; ++k) {
            auto it = edge.emplace(ekey(m.tris[t][k], m.tris[t][(k + 1) % 3]), std::array<int, 2>{-1, -1}).first;
            if (it->second[0] < 0) it->second[0] = int(t);
            else if (it->second[1] < 0) it->second[1] = int(t);
            else throw std::runtime_error("curved patches: an edge has more than two triangles");
        }
    m.ctri.clear();
    m.ctri.reserve(nt);
    for (size_t t = 0; t < nt; ++t) {
        const auto& tr = m.tris[t];
        Vec3 across[3];  // edges 1-2, 2-3, 3-1
        for (int k = 0; k < 3; ++k) {
            const auto& e = edge.at(ekey(tr[k], tr[(k + 1) % 3]));
            const int other = e[0] == int(t) ? e[1] : e[0];
            if (other < 0) throw std::runtime_error("curved patches need a closed surface");
            across[k] = tn[other];
        }
        m.ctri.emplace_back(m.verts[tr[0]], m.verts[tr[1]], m.verts[tr[2]], vn[tr[0]], vn[tr[1]], vn[tr[2]],
                            across[0], across[1], across[2]);
    }
}

void Mesh::patch_points(int vid, const Rule& rule, std::vector<QuadPt>& out) const {
    const double third = 1.0 / 3.0;
    const Vec3 centre_uvw(third, third, third);
    for (int kt = vt_ptr[vid]; kt < vt_ptr[vid + 1]; ++kt) {
        const int t = vt_tri[kt];
        const auto& tri = tris[t];
        const Vec3& v1 = verts[tri[0]];
        const Vec3& v2 = verts[tri[1]];
        const Vec3& v3 = verts[tri[2]];
        Vec3 cr = (v2 - v1).cross(v3 - v1);
        const double tri_area = 0.5 * cr.norm();
        const Vec3 tn = cr.unit();
        // barycentric (u, v, w) triples, point = w*v1 + u*v2 + v*v3
        Vec3 origin, mid1, mid2;
        if (tri[0] == vid) {
            origin = {0, 0, 1}; mid1 = {0.5, 0, 0.5}; mid2 = {0, 0.5, 0.5};
        } else if (tri[1] == vid) {
            origin = {1, 0, 0}; mid1 = {0.5, 0.5, 0}; mid2 = {0.5, 0, 0.5};
        } else {
            origin = {0, 1, 0}; mid1 = {0, 0.5, 0.5}; mid2 = {0.5, 0.5, 0};
        }
        const Vec3 sub[2][3] = {{centre_uvw, origin, mid1}, {mid2, origin, centre_uvw}};
        for (int s = 0; s < 2; ++s) {
            const Vec3& P1 = sub[s][0];
            const Vec3& P2 = sub[s][1];
            const Vec3& P3 = sub[s][2];
            for (size_t k = 0; k < rule.w.size(); ++k) {
                const double ru = rule.u[k], rv = rule.v[k];
                Vec3 uvw = P1 * (1.0 - ru - rv) + P2 * ru + P3 * rv;
                if (curved) {
                    // the parametric point mapped onto the curved patch; the
                    // weight takes the surface jacobian |dS/du x dS/dv|, which
                    // is twice the area for a flat triangle
                    double p[3], a[3], b[3];
                    beepcurve::eval(ctri[t].ctl, uvw.x, uvw.y, p, a, b);
                    const Vec3 du(a[0], a[1], a[2]), dv(b[0], b[1], b[2]);
                    const Vec3 cr2 = du.cross(dv);
                    const double jac = std::sqrt(du.dot(du) * dv.dot(dv) - du.dot(dv) * du.dot(dv));
                    out.push_back(QuadPt{Vec3(p[0], p[1], p[2]), cr2.unit(), rule.w[k] / 6.0 * jac / 2.0});
                    continue;
                }
                // each sub-triangle is one sixth of the triangle
                const double wt = rule.w[k] * tri_area / 6.0;
                Vec3 pos = v1 * uvw.z + v2 * uvw.x + v3 * uvw.y;
                out.push_back(QuadPt{pos, tn, wt});
            }
        }
    }
}

void Mesh::finalize() {
    const size_t nv = verts.size();
    vt_ptr.assign(nv + 1, 0);
    for (const auto& tr : tris)
        for (int k = 0; k < 3; ++k) vt_ptr[size_t(tr[k]) + 1]++;
    for (size_t v = 0; v < nv; ++v) vt_ptr[v + 1] += vt_ptr[v];
    vt_tri.resize(size_t(vt_ptr[nv]));
    vt_corner.resize(size_t(vt_ptr[nv]));
    {
        std::vector<int> at(vt_ptr.begin(), vt_ptr.end() - 1);
        for (size_t t = 0; t < tris.size(); ++t)
            for (int k = 0; k < 3; ++k) {
                const int q = at[size_t(tris[t][k])]++;
                vt_tri[q] = int(t);
                vt_corner[q] = k;
            }
    }
    if (curved) build_curved(*this);
    else ctri.clear();
    const Rule& g4 = rule_for_order(4);
    centroid.resize(verts.size());
    normal.resize(verts.size());
    area.resize(verts.size());
    std::vector<QuadPt> pts;
    for (size_t v = 0; v < verts.size(); ++v) {
        pts.clear();
        patch_points(int(v), g4, pts);
        Vec3 an, c;
        double a = 0;
        for (const auto& q : pts) {
            an += q.n * q.w;
            a += q.w;
            c += q.p * q.w;
        }
        normal[v] = an.unit();
        centroid[v] = c / a;
        area[v] = a;
    }
    finalize_charges();
}

void Mesh::finalize_charges() {
    Vec3 s;
    for (const auto& q : charges) s += Vec3(q[0], q[1], q[2]);
    if (!centre_given) centre = charges.empty() ? Vec3() : s / double(charges.size());
    radius = 0;
    for (const auto& c : centroid) radius = std::max(radius, (c - centre).norm());
}

Mesh make_icosphere(int subdiv, double radius) {
    Mesh m;
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    std::vector<Vec3> v = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0},
                           {0, -1, t}, {0, 1, t}, {0, -1, -t}, {0, 1, -t},
                           {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (auto& p : v) p = p.unit();
    std::vector<std::array<int, 3>> f = {
        {0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
        {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
        {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    for (int s = 0; s < subdiv; ++s) {
        std::unordered_map<uint64_t, int> mid;
        auto midpoint = [&](int a, int b) {
            uint64_t k = a < b ? (uint64_t(a) << 32 | uint32_t(b)) : (uint64_t(b) << 32 | uint32_t(a));
            auto it = mid.find(k);
            if (it != mid.end()) return it->second;
            v.push_back(((v[a] + v[b]) * 0.5).unit());
            int id = int(v.size()) - 1;
            mid.emplace(k, id);
            return id;
        };
        std::vector<std::array<int, 3>> nf;
        nf.reserve(f.size() * 4);
        for (const auto& tri : f) {
            int a = midpoint(tri[0], tri[1]);
            int b = midpoint(tri[1], tri[2]);
            int c = midpoint(tri[2], tri[0]);
            nf.push_back({tri[0], a, c});
            nf.push_back({tri[1], b, a});
            nf.push_back({tri[2], c, b});
            nf.push_back({a, b, c});
        }
        f.swap(nf);
    }
    for (auto& p : v) p = p * radius;
    m.verts = std::move(v);
    m.tris = std::move(f);
    return m;
}

// Real spherical harmonic Y_lm (orthonormal on the unit sphere).
static double real_ylm(int l, int m, double ct, double phi) {
    int am = std::abs(m);
    // associated Legendre P_l^am(ct) with Condon-Shortley phase
    double st = std::sqrt(std::max(0.0, 1.0 - ct * ct));
    double pmm = 1.0;
    for (int i = 1; i <= am; ++i) pmm *= -(2.0 * i - 1.0) * st;
    double plm;
    if (l == am) {
        plm = pmm;
    } else {
        double pmmp1 = ct * (2.0 * am + 1.0) * pmm;
        if (l == am + 1) {
            plm = pmmp1;
        } else {
            double pll = 0;
            for (int ll = am + 2; ll <= l; ++ll) {
                pll = ((2.0 * ll - 1.0) * ct * pmmp1 - (ll + am - 1.0) * pmm) / (ll - am);
                pmm = pmmp1;
                pmmp1 = pll;
            }
            plm = pll;
        }
    }
    double ratio = 1.0;  // (l-am)!/(l+am)!
    for (int i = l - am + 1; i <= l + am; ++i) ratio /= i;
    double norm = std::sqrt((2.0 * l + 1.0) / (4.0 * M_PI) * ratio);
    if (m == 0) return norm * plm;
    if (m > 0) return std::sqrt(2.0) * norm * plm * std::cos(am * phi);
    return std::sqrt(2.0) * norm * plm * std::sin(am * phi);
}

double StarShape::radius(const Vec3& dir) const {
    Vec3 d = dir.unit();
    double ct = d.z, phi = std::atan2(d.y, d.x);
    double s = 0;
    for (int l = 2; l <= lmax; ++l)
        for (int m = -l; m <= l; ++m) s += a[l * l + l + m] * real_ylm(l, m, ct, phi);
    return r0 * (1.0 + s);
}

StarShape star_shape(double r0, int lmax, double sigma0, uint64_t seed) {
    StarShape sh{r0, lmax, std::vector<double>((lmax + 1) * (lmax + 1), 0.0)};
    Rng rng(seed);
    for (int l = 2; l <= lmax; ++l)
        for (int m = -l; m <= l; ++m) sh.a[l * l + l + m] = rng.normal() * (sigma0 / l);
    return sh;
}

Mesh make_star(int subdiv, const StarShape& sh) {
    Mesh m = make_icosphere(subdiv, 1.0);
    for (auto& p : m.verts) p = p * sh.radius(p);
    return m;
}

void charges_uniform_ball(Mesh& m, int n, double rmax, double qmin, double qmax, uint64_t seed) {
    Rng rng(seed);
    m.charges.clear();
    while (int(m.charges.size()) < n) {
        Vec3 p(2 * rng.uniform() - 1, 2 * rng.uniform() - 1, 2 * rng.uniform() - 1);
        if (p.dot(p) >= 1.0) continue;
        p = p * rmax;
        double s = rng.uniform() < 0.5 ? -1.0 : 1.0;
        double q = s * (qmin + (qmax - qmin) * rng.uniform());
        m.charges.push_back({p.x, p.y, p.z, q});
    }
}

void charges_random_walk(Mesh& m, const StarShape& shape, int n, double step, double clearance,
                         double qsigma, uint64_t seed) {
    Rng rng(seed);
    m.charges.clear();
    Vec3 pos(0, 0, 0);
    for (int i = 0; i < n; ++i) {
        Vec3 next;
        bool ok = false;
        for (int attempt = 0; attempt < 200 && !ok; ++attempt) {
            Vec3 d(rng.normal(), rng.normal(), rng.normal());
            next = pos + d.unit() * step;
            double rn = next.norm();
            ok = rn < 1e-12 || rn <= shape.radius(next) - clearance;
        }
        if (!ok) next = pos * 0.5;  // step back toward the centre
        double q = rng.normal() * qsigma;
        m.charges.push_back({next.x, next.y, next.z, q});
        pos = next;
    }
}

}  // namespace beep
