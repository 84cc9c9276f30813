// bvh_build.hpp -- host BVH construction and the device layouts it feeds.
//
// Two device layouts are produced from one build:
//  * "reference order" FP64: the reference's DFS node array (left child at
//    node+1, right child at idx, leaf slots contiguous) used by the FP64
//    kernels and the classifier, so traversal order and leaf sets equal the
//    reference's (src/bvh.cpp:174-253, :295-387);
//  * "child-box" FP32: each internal node holds both child boxes (rounded
//    outward) and both child references in one 64 B (3D) / 48 B (2D) record,
//    so one load decides both children; leaves are 64 B (3D) / 16 B (2D)
//    slot records.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace wosx {

struct HostMesh {
    int dim = 0;
    std::vector<double> verts;      // n_verts x dim
    std::vector<int32_t> prims;     // n_prims x dim
    std::vector<int32_t> bids;      // n_prims (2D boundary ids; zeros in 3D)
    int64_t n_verts() const { return dim ? (int64_t)verts.size() / dim : 0; }
    int64_t n_prims() const { return dim ? (int64_t)prims.size() / dim : 0; }
};

// reference BVHNode (include/wos/bvh.hpp:10-15) with 3 coordinates
struct RefNode {
    double lo[3], hi[3];
    int32_t count;  // 0: internal
    int32_t idx;    // leaf: first slot; internal: right child
};

struct HostBVH {
    std::vector<RefNode> nodes;
    std::vector<int32_t> slots;  // leaf slot -> primitive id
    int depth = 0;               // nodes on the longest root-leaf path
};

// Build with the reference rule: leaf at <= 4 primitives, split on the
// longest box axis (ties x, then y), nth_element at (start+end)/2 by centroid
// with the primitive index as tie-break.  Returns an empty BVH on error.
HostBVH build_reference_bvh(const HostMesh &m, std::string *err);

// ---------------- device records ----------------
// FP64 reference-order nodes
struct alignas(8) NodeD2 { double lox, loy, hix, hiy; int32_t count, idx; };
struct alignas(8) NodeD3 { double lox, loy, loz, hix, hiy, hiz; int32_t count, idx; };
// FP64 leaf slots
struct alignas(16) SegD { double s0x, s0y, s1x, s1y; };
// Eberly cache of src/bvh.cpp:190-205 (v0, edge1, edge2, a, b, c, det, denom)
struct alignas(16) TriD {
    double v0x, v0y, v0z, e1x, e1y, e1z, e2x, e2y, e2z, a, b, c, det, denom, pad0, pad1;
};

// FP32 child-box nodes.  Child reference: >= 0 internal node index,
// < 0 leaf = -(1 + (first_slot << 3 | count)).
struct alignas(16) NodeF3 {
    float l_lox, l_loy, l_loz, r_lox;
    float r_loy, r_loz, l_hix, l_hiy;
    float l_hiz, r_hix, r_hiy, r_hiz;
    int32_t left, right, pad0, pad1;
};
struct alignas(16) NodeF2 {
    float l_lox, l_loy, l_hix, l_hiy;
    float r_lox, r_loy, r_hix, r_hiy;
    int32_t left, right, pad0, pad1;
};
// FP32 4-wide nodes: the binary FP32 tree with every second level folded in
// (a node's children are its binary children, or their children where those
// are internal).  Up to four children's boxes (rounded outward, each
// coordinate a row of four) and references; an unused child has the empty
// box lo = +inf, hi = -inf, which is never entered.
struct alignas(16) NodeF3W {
    float lox[4], loy[4], loz[4], hix[4], hiy[4], hiz[4];
    int32_t ref[4];
    int32_t pad[4];
};
struct alignas(16) NodeF2W {
    float lox[4], loy[4], hix[4], hiy[4];
    int32_t ref[4];
};
// FP32 leaf slots.  A triangle (48 B: three 16 B loads) in its own frame,
// formed in FP64: origin o at one end of its longest edge, unit vector u
// along that edge, unit vector w in the plane towards the third vertex
// (n = u x w is the normal); the triangle is (0, 0), (l1, 0), (x2, y2) in
// (u, w) coordinates with y2 >= 0.  Degenerate triangles keep an orthonormal
// frame (y2 = 0, or l1 = x2 = y2 = 0 for a point).
struct alignas(16) TriF {
    float ox, oy, oz, l1;
    float ux, uy, uz, x2;
    float wx, wy, wz, y2;
};
struct alignas(16) SegF { float s0x, s0y, vx, vy; };

// One domain of a forest: its slice of the concatenated arrays.
struct DomainInfo {
    int64_t vert_off = 0, n_verts = 0;
    int64_t prim_off = 0, n_prims = 0;   // primitives == leaf slots
    int32_t node_off = 0;                // FP64 root node (reference order)
    int32_t root_ref = 0;                // FP32 root reference
    int32_t root_ref_w = 0;              // FP32 4-wide root reference
    int depth = 0;
    double bounds[6] = {0, 0, 0, 0, 0, 0};  // bbox of the domain's vertices (mesh.cpp:296-322)
    double coord_scale = 1.0;               // max |coordinate| of this domain, >= 1
};

struct DeviceLayouts {
    std::vector<NodeD2> nd2;
    std::vector<NodeD3> nd3;
    std::vector<SegD> segd;     // slot order
    std::vector<TriD> trid;     // slot order
    std::vector<NodeF2> nf2;
    std::vector<NodeF3> nf3;
    std::vector<NodeF2W> nw2;   // the 4-wide forest (breadth-first)
    std::vector<NodeF3W> nw3;
    std::vector<SegF> segf;
    std::vector<TriF> trif;
    std::vector<int32_t> slot_prim; // global slot -> global primitive
    std::vector<DomainInfo> domains;
    int depth = 0;             // deepest domain
    int depth_w = 0;           // deepest domain of the 4-wide forest (internal levels)
    double coord_scale = 1.0;  // max |coordinate| over all domains, >= 1
};

// Concatenate domains (each with its own reference-rule BVH) into one set of
// device arrays.  Vertex, primitive, slot and node indices become global; the
// FP32 internal nodes are numbered breadth-first over the whole forest, so
// the top levels of every domain form the prefix staged in LDS.
DeviceLayouts make_layouts(const std::vector<const HostMesh *> &meshes,
                           const std::vector<const HostBVH *> &bvhs);

}  // namespace wosx
