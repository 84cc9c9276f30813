// obj_mesh.cpp -- OBJ reader of the `wos` command line tool (behaviour of
// src/mesh.cpp:51-331, written for this tool).
#include "obj_mesh.hpp"

#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>

namespace woscli {

namespace {

[[noreturn]] void mesh_error(const char *path, int line, const std::string &msg) {
    std::string where = std::string("Mesh ") + path;
    if (line > 0) where += ":" + std::to_string(line);
    throw std::runtime_error(where + ": " + msg);
}

// 1-based reference "i", "i/t" or "i/t/n" -> 0-based index
int vertex_ref(const std::string &token, int total, const char *path, int line) {
    const std::string idx = token.substr(0, token.find('/'));
    if (idx.empty()) mesh_error(path, line, "missing vertex index in '" + token + "'");
    char *end = nullptr;
    errno = 0;
    const long v = std::strtol(idx.c_str(), &end, 10);
    if (errno == ERANGE || end == idx.c_str() || *end != '\0' || v <= 0 || v > INT_MAX)
        mesh_error(path, line, "vertex index must be a positive 1-based integer: '" + token + "'");
    if (v - 1 >= total) mesh_error(path, line, "vertex index out of range: '" + token + "'");
    return (int)(v - 1);
}

bool next_kind(std::istringstream &rec, std::string &kind) {
    return (rec >> kind) && kind[0] != '#';
}

}  // namespace

int peek_mesh_dim(const char *path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error(std::string("Could not read mesh file ") + path);
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream rec(line);
        std::string kind;
        if (!next_kind(rec, kind)) continue;
        if (kind == "l") return 2;
        if (kind == "f") return 3;
    }
    throw std::runtime_error(std::string("Could not determine mesh dimension in ") + path +
                             " (no 'l' or 'f' entries)");
}

ObjMesh load_mesh(const char *path, int dim) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error(std::string("Could not read mesh file ") + path);
    std::string line;
    int n_v = 0, n_p = 0, ln = 0;
    // first pass: counts and record shapes
    while (std::getline(in, line)) {
        ++ln;
        std::istringstream rec(line);
        std::string kind;
        if (!next_kind(rec, kind)) continue;
        if (kind == "v") {
            ++n_v;
            continue;
        }
        if ((dim == 2 && kind == "l") || (dim == 3 && kind == "f")) {
            std::string tok;
            int k = 0;
            while (rec >> tok && tok[0] != '#') ++k;
            if (dim == 2) {
                if (k < 2) mesh_error(path, ln, "polyline has fewer than 2 vertices");
                n_p += k;  // every consecutive pair plus a possible closing segment
            } else {
                if (k != 3)
                    mesh_error(path, ln, "face has " + std::to_string(k) +
                                             " vertices; only triangles are supported");
                ++n_p;
            }
        }
    }
    if (n_v == 0) mesh_error(path, 0, "contains no vertices");
    if (n_p == 0)
        mesh_error(path, 0, dim == 2 ? "contains no line primitives" : "contains no triangle primitives");
    in.clear();
    in.seekg(0, std::ios::beg);
    if (!in) throw std::runtime_error(std::string("Could not rewind mesh file ") + path);

    ObjMesh m;
    m.dim = dim;
    m.verts.reserve((size_t)n_v * dim);
    m.prims.reserve((size_t)n_p * dim);
    int boundary = 0;
    ln = 0;
    while (std::getline(in, line)) {
        ++ln;
        std::istringstream rec(line);
        std::string kind;
        if (!next_kind(rec, kind)) continue;
        if (kind == "v") {
            double x, y, z = 0.0;
            if (!(rec >> x >> y)) mesh_error(path, ln, "malformed vertex record");
            std::string zt;
            if (rec >> zt && zt[0] != '#') {
                char *end = nullptr;
                errno = 0;
                z = std::strtod(zt.c_str(), &end);
                if (errno == ERANGE || end == zt.c_str() || *end != '\0')
                    mesh_error(path, ln, "malformed vertex record");
            }
            if (!std::isfinite(x) || !std::isfinite(y) || !std::isfinite(z))
                mesh_error(path, ln, "vertex coordinates must be finite");
            m.verts.push_back(x);
            m.verts.push_back(y);
            if (dim == 3) m.verts.push_back(z);
        } else if (dim == 2 && kind == "l") {
            int first = -1, prev = -1;
            std::string tok;
            while (rec >> tok) {
                if (tok[0] == '#') break;
                const int v = vertex_ref(tok, n_v, path, ln);
                if (first < 0) first = v;
                if (prev >= 0) {
                    m.prims.push_back(prev);
                    m.prims.push_back(v);
                    m.bids.push_back(boundary);
                }
                prev = v;
            }
            if (prev >= 0 && prev != first) {
                std::fprintf(stderr, "Warning: open polyline in %s auto-closed\n", path);
                m.prims.push_back(prev);
                m.prims.push_back(first);
                m.bids.push_back(boundary);
            }
            ++boundary;
        } else if (dim == 3 && kind == "f") {
            std::string refs[3];
            for (auto &r : refs)
                if (!(rec >> r) || r[0] == '#') mesh_error(path, ln, "malformed triangle record");
            for (auto &r : refs) m.prims.push_back(vertex_ref(r, n_v, path, ln));
        }
    }
    const double *V = m.verts.data();
    for (int64_t p = 0; p < m.n_prims(); ++p) {
        if (dim == 2) {
            const double *a = V + 2 * m.prims[2 * p], *b = V + 2 * m.prims[2 * p + 1];
            const double dx = b[0] - a[0], dy = b[1] - a[1];
            if (!(dx * dx + dy * dy > 0.0)) mesh_error(path, 0, "contains a zero-length line segment");
        } else {
            const double *a = V + 3 * m.prims[3 * p], *b = V + 3 * m.prims[3 * p + 1],
                         *c = V + 3 * m.prims[3 * p + 2];
            const double ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
            const double ac[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
            const double cx = ab[1] * ac[2] - ab[2] * ac[1];
            const double cy = ab[2] * ac[0] - ab[0] * ac[2];
            const double cz = ab[0] * ac[1] - ab[1] * ac[0];
            if (!(cx * cx + cy * cy + cz * cz > 0.0)) mesh_error(path, 0, "contains a degenerate triangle");
        }
    }
    return m;
}

}  // namespace woscli
