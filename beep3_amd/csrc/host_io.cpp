// Mesh and charge file formats (host C++): OFF and GTS surfaces
// (src/bem/off_utils.h:26-160, src/bem/gts_utils.h:32-140) and x y z q r
// charge lists (the xyzqr member of the reference's .mtz tarball,
// src/bem/mesh_tarball.h:52-75).  Comment lines start with '#' or '!'.
#include <array>
#include <fstream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>

#include "../../include/beep_host.h"
#include "host.hpp"

using namespace beep;

namespace {

std::string next_line(std::istream& in) {
    std::string line;
    while (std::getline(in, line)) {
        size_t p = line.find_first_not_of(" \t\r");
        if (p == std::string::npos) continue;
        if (line[p] == '#' || line[p] == '!') continue;
        return line;
    }
    throw std::runtime_error("unexpected end of file");
}

void read_off(std::istream& in, Mesh& m) {
    std::string head = next_line(in);
    long nv = 0, nf = 0, ne = 0;
    std::stringstream hs(head);
    std::string kw;
    hs >> kw;
    if (kw == "OFF" || kw == "nOFF") {
        std::stringstream cs(next_line(in));
        cs >> nv >> nf >> ne;
    } else {
        std::stringstream cs(head);
        if (!(cs >> nv >> nf >> ne)) throw std::runtime_error("OFF: bad header");
    }
    if (nv <= 0 || nf <= 0) throw std::runtime_error("OFF: empty mesh");
    for (long i = 0; i < nv; ++i) {
        std::stringstream ls(next_line(in));
        double x, y, z;
        if (!(ls >> x >> y >> z)) throw std::runtime_error("OFF: bad vertex line");
        m.verts.push_back(Vec3(x, y, z));
    }
    for (long f = 0; f < nf; ++f) {
        std::stringstream ls(next_line(in));
        int k, a, b, c;
        if (!(ls >> k >> a >> b >> c) || k != 3)
            throw std::runtime_error("OFF: faces must be triangles");
        if (a < 0 || b < 0 || c < 0 || a >= nv || b >= nv || c >= nv)
            throw std::runtime_error("OFF: vertex index out of range");
        m.tris.push_back({a, b, c});
    }
}

void read_gts(std::istream& in, Mesh& m) {
    std::stringstream hs(next_line(in));
    long nv = 0, ne = 0, nf = 0;
    if (!(hs >> nv >> ne >> nf) || nv <= 0 || nf <= 0) throw std::runtime_error("GTS: bad header");
    for (long i = 0; i < nv; ++i) {
        std::stringstream ls(next_line(in));
        double x, y, z;
        if (!(ls >> x >> y >> z)) throw std::runtime_error("GTS: bad vertex line");
        m.verts.push_back(Vec3(x, y, z));
    }
    std::vector<std::array<int, 2>> edges(ne);
    for (long e = 0; e < ne; ++e) {
        std::stringstream ls(next_line(in));
        int a, b;
        if (!(ls >> a >> b) || a < 1 || b < 1 || a > nv || b > nv)
            throw std::runtime_error("GTS: bad edge line");
        edges[e] = {a - 1, b - 1};
    }
    for (long f = 0; f < nf; ++f) {
        std::stringstream ls(next_line(in));
        int e[3];
        if (!(ls >> e[0] >> e[1] >> e[2])) throw std::runtime_error("GTS: bad face line");
        std::array<int, 3> tri{};
        for (int i = 0; i < 3; ++i) {
            const int ea = e[i] - 1, eb = e[(i + 1) % 3] - 1;
            if (ea < 0 || eb < 0 || ea >= ne || eb >= ne) throw std::runtime_error("GTS: edge index");
            const auto& A = edges[ea];
            const auto& B = edges[eb];
            // the vertex shared by consecutive edges (gts_utils.h:99-120)
            int common = -1;
            for (int p : A)
                for (int q : B)
                    if (p == q) common = p;
            if (common < 0) throw std::runtime_error("GTS: face edges do not form a triangle");
            tri[i] = common;
        }
        m.tris.push_back(tri);
    }
}

}  // namespace

extern "C" int beep_mesh_read(const char* path, beep_mesh** out) {
    try {
        if (!path || !out) throw std::runtime_error("null argument");
        std::ifstream in(path);
        if (!in.good()) throw std::runtime_error(std::string("cannot open ") + path);
        auto m = std::make_unique<beep_mesh>();
        const std::string p(path);
        if (p.size() >= 4 && p.substr(p.size() - 4) == ".off") read_off(in, m->mesh);
        else if (p.size() >= 4 && p.substr(p.size() - 4) == ".gts") read_gts(in, m->mesh);
        else throw std::runtime_error("mesh file must end in .off or .gts");
        m->mesh.finalize();
        *out = m.release();
        return 0;
    } catch (const std::exception& e) {
        beep_host_set_error(e.what());
        return -1;
    }
}

extern "C" int beep_mesh_write_off(const beep_mesh* m, const char* path) {
    try {
        if (!m || !path) throw std::runtime_error("null argument");
        std::ofstream out(path);
        if (!out.good()) throw std::runtime_error(std::string("cannot write ") + path);
        out.precision(17);
        out << "OFF\n" << m->mesh.verts.size() << " " << m->mesh.tris.size() << " 0\n";
        for (const auto& v : m->mesh.verts) out << v.x << " " << v.y << " " << v.z << "\n";
        for (const auto& t : m->mesh.tris) out << "3 " << t[0] << " " << t[1] << " " << t[2] << "\n";
        return out.good() ? 0 : (beep_host_set_error("write failed"), -1);
    } catch (const std::exception& e) {
        beep_host_set_error(e.what());
        return -1;
    }
}

extern "C" int beep_mesh_read_charges(beep_mesh* m, const char* path) {
    try {
        if (!m || !path) throw std::runtime_error("null argument");
        std::ifstream in(path);
        if (!in.good()) throw std::runtime_error(std::string("cannot open ") + path);
        m->mesh.charges.clear();
        std::string line;
        while (std::getline(in, line)) {
            size_t p = line.find_first_not_of(" \t\r");
            if (p == std::string::npos || line[p] == '#' || line[p] == '!') continue;
            std::stringstream ls(line);
            double x, y, z, q;
            if (!(ls >> x >> y >> z >> q)) throw std::runtime_error("charges: need x y z q [r]");
            m->mesh.charges.push_back({x, y, z, q});
        }
        m->mesh.finalize_charges();
        m->stamp = next_mesh_stamp();
        return 0;
    } catch (const std::exception& e) {
        beep_host_set_error(e.what());
        return -1;
    }
}
