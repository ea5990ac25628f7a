// Mesh and charge file formats (host C++): OFF and GTS surfaces
// (src/bem/off_utils.h:26-160, src/bem/gts_utils.h:32-140) and x y z q r
// charge lists (the xyzqr member of the reference's .mtz tarball,
// src/bem/mesh_tarball.h:52-75), and the .mtz tarball itself (a gzip or
// plain tar archive described by its definition.xml, mesh_tarball.cpp:
// 118-270, read as Mesh::init(mtz) does, mesh.cpp:88-151, for planar
// patches).  Comment lines start with '#' or '!'.
#include <zlib.h>

#include <array>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <vector>
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

void read_charges(std::istream& in, Mesh& m) {
    m.charges.clear();
    std::string line;
    while (std::getline(in, line)) {
        size_t p = line.find_first_not_of(" \t\r");
        if (p == std::string::npos || line[p] == '#' || line[p] == '!') continue;
        std::stringstream ls(line);
        double x, y, z, q;
        if (!(ls >> x >> y >> z >> q)) throw std::runtime_error("charges: need x y z q [r]");
        m.charges.push_back({x, y, z, q});
    }
}

// ---- .mtz: a tar archive, gzip-compressed or plain, whose definition.xml
// names its members (MeshTarball, mesh_tarball.cpp:118-270) ----

// The whole file, inflated if it is gzip (zlib reads a plain file as it is).
std::string read_maybe_gzip(const char* path) {
    {
        std::ifstream raw(path, std::ios::binary);
        if (!raw.good()) throw std::runtime_error(std::string("cannot open ") + path);
        unsigned char magic[3] = {0, 0, 0};
        raw.read(reinterpret_cast<char*>(magic), 3);
        if (magic[0] == 'B' && magic[1] == 'Z' && magic[2] == 'h')
            throw std::runtime_error(".mtz: bzip2 compression is not read here; recompress with gzip");
        if (magic[0] == 0x1f && magic[1] == 0x9d)
            throw std::runtime_error(".mtz: compress (.Z) compression is not read here; recompress with gzip");
    }
    gzFile f = gzopen(path, "rb");
    if (!f) throw std::runtime_error(std::string("cannot open ") + path);
    std::string out;
    std::vector<char> buf(size_t(1) << 16);
    int n = 0;
    while ((n = gzread(f, buf.data(), unsigned(buf.size()))) > 0) out.append(buf.data(), size_t(n));
    int err = Z_OK;
    const std::string msg = n < 0 ? gzerror(f, &err) : "";
    gzclose(f);
    if (n < 0) throw std::runtime_error(".mtz: " + msg);
    return out;
}

// an octal (or base-256) numeric field of a tar header
uint64_t tar_number(const char* f, size_t len) {
    if (static_cast<unsigned char>(f[0]) & 0x80) {
        uint64_t v = static_cast<unsigned char>(f[0]) & 0x7f;
        for (size_t k = 1; k < len; ++k) v = (v << 8) | static_cast<unsigned char>(f[k]);
        return v;
    }
    uint64_t v = 0;
    size_t k = 0;
    while (k < len && (f[k] == ' ' || f[k] == '\0')) ++k;
    for (; k < len && f[k] >= '0' && f[k] <= '7'; ++k) v = v * 8 + uint64_t(f[k] - '0');
    return v;
}

std::string strip_dot_slash(std::string n) {
    while (n.size() >= 2 && n[0] == '.' && n[1] == '/') n.erase(0, 2);
    return n;
}

// Regular files of a tar archive by path (ustar names with their prefix, GNU
// long names and pax path records), headers checked against their checksums.
std::map<std::string, std::string> untar(const std::string& t) {
    std::map<std::string, std::string> files;
    std::string next_name;
    size_t off = 0;
    while (off + 512 <= t.size()) {
        const char* h = t.data() + off;
        bool zero = true;
        for (int k = 0; k < 512 && zero; ++k) zero = h[k] == 0;
        if (zero) break;  // end-of-archive block
        uint64_t sum = 0;
        for (int k = 0; k < 512; ++k)
            sum += (k >= 148 && k < 156) ? uint64_t(' ') : uint64_t(static_cast<unsigned char>(h[k]));
        if (sum != tar_number(h + 148, 8)) throw std::runtime_error(".mtz: not a tar archive (bad header checksum)");
        const uint64_t size = tar_number(h + 124, 12);
        const char type = h[156];
        off += 512;
        if (size > t.size() - off) throw std::runtime_error(".mtz: truncated archive");
        const std::string body = t.substr(off, size_t(size));
        off += size_t((size + 511) / 512 * 512);
        std::string name(h, strnlen(h, 100));
        if (std::memcmp(h + 257, "ustar", 5) == 0 && h[345]) name = std::string(h + 345, strnlen(h + 345, 155)) + "/" + name;
        if (type == 'L') {  // GNU long name of the next member
            next_name = std::string(body.c_str());
            continue;
        }
        if (type == 'x') {  // pax records "len key=value\n"; keep path=
            size_t p = 0;
            while (p < body.size()) {
                const size_t sp = body.find(' ', p);
                if (sp == std::string::npos) break;
                const size_t len = size_t(std::strtoul(body.c_str() + p, nullptr, 10));
                if (len == 0 || p + len > body.size()) break;
                const std::string rec = body.substr(sp + 1, p + len - sp - 2);
                if (rec.compare(0, 5, "path=") == 0) next_name = rec.substr(5);
                p += len;
            }
            continue;
        }
        if (!next_name.empty()) {
            name = next_name;
            next_name.clear();
        }
        if (type == '0' || type == '\0' || type == '7') files[strip_dot_slash(name)] = body;
    }
    return files;
}

std::string trim(const std::string& v) {
    const size_t a = v.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    return v.substr(a, v.find_last_not_of(" \t\r\n") - a + 1);
}

// The text of the <tag> element under the root element of definition.xml
// ("" if absent); two such elements are an error (mesh_tarball.cpp:205-265).
std::string xml_child_text(const std::string& body, const std::string& tag) {
    std::string text;
    int seen = 0;
    size_t p = 0;
    while ((p = body.find("<" + tag, p)) != std::string::npos) {
        const char c = p + 1 + tag.size() < body.size() ? body[p + 1 + tag.size()] : '\0';
        if (c != '>' && c != ' ' && c != '\t' && c != '\n' && c != '\r' && c != '/') {
            p += 1;
            continue;
        }
        const size_t gt = body.find('>', p);
        if (gt == std::string::npos) throw std::runtime_error("definition.xml: unterminated <" + tag + ">");
        ++seen;
        if (body[gt - 1] == '/') {  // <tag/>: empty
            text.clear();
        } else {
            const size_t end = body.find("</" + tag, gt);
            if (end == std::string::npos) throw std::runtime_error("definition.xml: no </" + tag + ">");
            text = trim(body.substr(gt + 1, end - gt - 1));
        }
        p = gt + 1;
    }
    if (seen > 1) throw std::runtime_error("definition.xml: multiple <" + tag + "> elements");
    return text;
}

// definition.xml without its declaration and comments, checked for the root
// element <contents>; returns the root's body
std::string xml_root_body(std::string x) {
    for (const char* open : {"<!--", "<?"}) {
        const std::string close = std::string(open) == "<!--" ? "-->" : "?>";
        size_t a;
        while ((a = x.find(open)) != std::string::npos) {
            const size_t b = x.find(close, a);
            if (b == std::string::npos) throw std::runtime_error("definition.xml: unterminated markup");
            x.erase(a, b + close.size() - a);
        }
    }
    const size_t a = x.find('<');
    if (a == std::string::npos) throw std::runtime_error("definition.xml: no root element");
    const std::string root = "contents";
    if (x.compare(a + 1, root.size(), root) != 0)
        throw std::runtime_error("definition.xml: the root element must be <contents>");
    const size_t gt = x.find('>', a);
    const size_t end = x.rfind("</" + root);
    if (gt == std::string::npos || end == std::string::npos || end < gt)
        throw std::runtime_error("definition.xml: unterminated <contents>");
    return x.substr(gt + 1, end - gt - 1);
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

extern "C" int beep_mesh_read_mtz(const char* path, beep_mesh** out) {
    try {
        if (!path || !out) throw std::runtime_error("null argument");
        const auto files = untar(read_maybe_gzip(path));
        auto member = [&](const std::string& name, const char* what) -> const std::string& {
            auto it = files.find(strip_dot_slash(name));
            if (it == files.end())
                throw std::runtime_error(std::string(".mtz: the ") + what + " member '" + name + "' is missing");
            return it->second;
        };
        const std::string root = xml_root_body(member("definition.xml", "definition"));
        // mandatory tags (mesh_tarball.cpp:208-223); energies holds the
        // curved-patch energy precalcs, recomputed for planar patches
        std::string mesh_name, xyzqr_name, centre_name;
        for (const char* tag : {"xyzqr", "mesh", "energies", "centre"}) {
            const std::string v = xml_child_text(root, tag);
            if (v.empty()) throw std::runtime_error(std::string("definition.xml: missing <") + tag + ">");
            if (std::string(tag) == "mesh") mesh_name = v;
            if (std::string(tag) == "xyzqr") xyzqr_name = v;
            if (std::string(tag) == "centre") centre_name = v;
        }
        const std::string fh_name = xml_child_text(root, "fh");
        for (const char* tag : {"pdb", "pqr", "vertex_normals", "diffusion_tensor", "ellipsoid", "ecm",
                                "mesh2", "energies2"})
            (void)xml_child_text(root, tag);  // optional; checked for duplicates only
        auto m = std::make_unique<beep_mesh>();
        {
            std::istringstream in(member(mesh_name, "mesh"));
            const std::string ext = mesh_name.size() >= 4 ? mesh_name.substr(mesh_name.size() - 4) : "";
            if (ext == ".gts") read_gts(in, m->mesh);
            else read_off(in, m->mesh);  // the reference's default surface format
        }
        {
            std::istringstream in(member(xyzqr_name, "xyzqr"));
            read_charges(in, m->mesh);
        }
        {
            std::istringstream in(member(centre_name, "centre"));
            double x, y, z;
            if (!(in >> x >> y >> z)) throw std::runtime_error(".mtz: the centre member needs x y z");
            m->mesh.centre = Vec3(x, y, z);
            m->mesh.centre_given = true;
        }
        m->mesh.finalize();
        if (!fh_name.empty()) {
            std::istringstream in(member(fh_name, "fh"));
            const size_t nv = m->mesh.verts.size();
            m->mesh.init_f.resize(nv);
            m->mesh.init_h.resize(nv);
            for (size_t k = 0; k < nv; ++k)
                if (!(in >> m->mesh.init_f[k] >> m->mesh.init_h[k]))
                    throw std::runtime_error(".mtz: the fh member needs f h for every node patch");
        }
        *out = m.release();
        return 0;
    } catch (const std::exception& e) {
        beep_host_set_error(e.what());
        return -1;
    }
}

extern "C" int beep_mesh_initial_fh(const beep_mesh* m, double* f, double* h) {
    try {
        if (!m || !f || !h) throw std::runtime_error("null argument");
        if (m->mesh.init_f.empty()) return 1;
        std::copy(m->mesh.init_f.begin(), m->mesh.init_f.end(), f);
        std::copy(m->mesh.init_h.begin(), m->mesh.init_h.end(), h);
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
        read_charges(in, m->mesh);
        m->mesh.finalize_charges();
        m->stamp = next_mesh_stamp();
        return 0;
    } catch (const std::exception& e) {
        beep_host_set_error(e.what());
        return -1;
    }
}
