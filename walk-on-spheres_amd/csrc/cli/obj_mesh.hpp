// obj_mesh.hpp -- OBJ reader of the `wos` command line tool.
//
// Same records, validation and messages as the reference loader
// (src/mesh.cpp:51-331): `v x y [z]`, 2D `l` polylines (each record is one
// boundary id; an open polyline is closed with a warning), 3D `f` triangles,
// 1-based vertex references, finite coordinates, no zero-length segments or
// degenerate triangles.  Errors are std::runtime_error with the reference's
// text.
#pragma once
#include <cstdint>
#include <vector>

namespace woscli {

struct ObjMesh {
    int dim = 0;
    std::vector<double> verts;    // n_verts x dim
    std::vector<int32_t> prims;   // n_prims x dim (0-based)
    std::vector<int32_t> bids;    // 2D boundary id per segment
    int64_t n_verts() const { return dim ? (int64_t)verts.size() / dim : 0; }
    int64_t n_prims() const { return dim ? (int64_t)prims.size() / dim : 0; }
};

// dimension of the first `l` (2) or `f` (3) record (src/mesh.cpp:51-74)
int peek_mesh_dim(const char *path);

// two-pass load (src/mesh.cpp:76-270)
ObjMesh load_mesh(const char *path, int dim);

}  // namespace woscli
