// bvh_build.cpp -- host BVH construction (reference split rule) and the
// flattening into the FP64 reference-order and FP32 child-box layouts.
#include "bvh_build.hpp"

#include <algorithm>
#include <cmath>
#include <limits>

namespace wosx {

namespace {

struct Item {
    double lo[3], hi[3], c[3];
    int32_t prim;
};

struct Builder {
    const HostMesh &m;
    std::vector<Item> items;
    HostBVH out;
    int node_ptr = 0;

    explicit Builder(const HostMesh &mesh) : m(mesh) {}

    // src/bvh.cpp:174-230: union box, leaf at <= 4, longest axis,
    // median split with the (centroid, primitive) order
    int build(int start, int end, int level) {
        const int dim = m.dim;
        const int span = end - start;
        const int node = node_ptr++;
        RefNode &nd = out.nodes[node];
        for (int k = 0; k < 3; ++k) { nd.lo[k] = items[start].lo[k]; nd.hi[k] = items[start].hi[k]; }
        for (int i = start + 1; i < end; ++i)
            for (int k = 0; k < dim; ++k) {
                nd.lo[k] = nd.lo[k] < items[i].lo[k] ? nd.lo[k] : items[i].lo[k];
                nd.hi[k] = nd.hi[k] > items[i].hi[k] ? nd.hi[k] : items[i].hi[k];
            }
        out.depth = std::max(out.depth, level);
        if (span <= 4) {
            nd.count = span;
            nd.idx = start;
            for (int i = 0; i < span; ++i) out.slots[start + i] = items[start + i].prim;
            return node;
        }
        nd.count = 0;
        const double ex = nd.hi[0] - nd.lo[0], ey = nd.hi[1] - nd.lo[1];
        int axis;
        if (dim == 2) {
            axis = (ex >= ey) ? 0 : 1;
        } else {
            const double ez = nd.hi[2] - nd.lo[2];
            axis = (ex >= ey) ? (ex >= ez ? 0 : 2) : (ey >= ez ? 1 : 2);
        }
        const int mid = (start + end) / 2;
        std::nth_element(items.begin() + start, items.begin() + mid, items.begin() + end,
                         [axis](const Item &a, const Item &b) {
                             if (a.c[axis] < b.c[axis]) return true;
                             if (a.c[axis] > b.c[axis]) return false;
                             return a.prim < b.prim;
                         });
        build(start, mid, level + 1);
        const int right = build(mid, end, level + 1);
        out.nodes[node].idx = right;
        return node;
    }
};

float round_down(double v) {
    float f = (float)v;
    if ((double)f > v) f = std::nextafter(f, -INFINITY);
    return f;
}
float round_up(double v) {
    float f = (float)v;
    if ((double)f < v) f = std::nextafter(f, INFINITY);
    return f;
}

}  // namespace

HostBVH build_reference_bvh(const HostMesh &m, std::string *err) {
    Builder b(m);
    const int64_t np = m.n_prims();
    if (np <= 0 || np > (1 << 27)) {
        if (err) *err = "mesh must have between 1 and 2^27 primitives";
        return HostBVH{};
    }
    b.items.resize(np);
    const int dim = m.dim;
    for (int64_t p = 0; p < np; ++p) {
        Item &it = b.items[p];
        it.prim = (int32_t)p;
        const double *v0 = &m.verts[(size_t)m.prims[dim * p] * dim];
        const double *v1 = &m.verts[(size_t)m.prims[dim * p + 1] * dim];
        for (int k = 0; k < 3; ++k) { it.lo[k] = it.hi[k] = it.c[k] = 0.0; }
        if (dim == 2) {
            // prim_bbox / centroid of src/mesh.cpp:324-366
            for (int k = 0; k < 2; ++k) {
                it.lo[k] = std::fmin(v0[k], v1[k]);
                it.hi[k] = std::fmax(v0[k], v1[k]);
                it.c[k] = (v0[k] + v1[k]) / 2.0;
            }
        } else {
            const double *v2 = &m.verts[(size_t)m.prims[dim * p + 2] * dim];
            for (int k = 0; k < 3; ++k) {
                it.lo[k] = std::fmin(std::fmin(v0[k], v1[k]), v2[k]);
                it.hi[k] = std::fmax(std::fmax(v0[k], v1[k]), v2[k]);
                it.c[k] = (v0[k] + v1[k] + v2[k]) / 3.0;
            }
        }
    }
    b.out.nodes.resize(2 * np - 1);
    b.out.slots.resize(np);
    b.build(0, (int)np, 1);
    b.out.nodes.resize(b.node_ptr);
    return std::move(b.out);
}

// The FP32 frame record of triangle (i0, i1, i2) of m, in FP64.
static TriF triangle_frame(const HostMesh &m, int32_t i0, int32_t i1, int32_t i2) {
    const int32_t id[3] = {i0, i1, i2};
    double P[3][3];
    for (int k = 0; k < 3; ++k)
        for (int j = 0; j < 3; ++j) P[k][j] = m.verts[(size_t)id[k] * 3 + j];
    auto dot = [](const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
    // the longest edge A -> B, C the third vertex
    int ia = 0;
    double best = -1.0;
    for (int k = 0; k < 3; ++k) {
        const double *a = P[k], *b = P[(k + 1) % 3];
        const double d[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        const double l = dot(d, d);
        if (l > best) { best = l; ia = k; }
    }
    const double *A = P[ia], *B = P[(ia + 1) % 3], *C = P[(ia + 2) % 3];
    double u[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
    const double l1 = std::sqrt(dot(u, u));
    if (l1 > 0.0) for (double &x : u) x /= l1;
    else { u[0] = 1.0; u[1] = 0.0; u[2] = 0.0; }
    const double e[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
    const double x2 = l1 > 0.0 ? dot(e, u) : 0.0;
    double w[3] = {e[0] - x2 * u[0], e[1] - x2 * u[1], e[2] - x2 * u[2]};
    // once more against u: w stays orthogonal to u when it is much shorter than e
    const double k2 = dot(w, u);
    for (int j = 0; j < 3; ++j) w[j] -= k2 * u[j];
    double y2 = std::sqrt(dot(w, w));
    if (l1 > 0.0 && y2 > 1e-9 * l1) {
        for (double &x : w) x /= y2;
    } else {
        // any unit vector orthogonal to u
        y2 = 0.0;
        const double ax[3] = {std::fabs(u[0]) < 0.9 ? 1.0 : 0.0, std::fabs(u[0]) < 0.9 ? 0.0 : 1.0, 0.0};
        const double k = dot(ax, u);
        double v[3] = {ax[0] - k * u[0], ax[1] - k * u[1], ax[2] - k * u[2]};
        const double lv = std::sqrt(dot(v, v));
        for (int j = 0; j < 3; ++j) w[j] = v[j] / lv;
    }
    TriF f;
    f.ox = (float)A[0]; f.oy = (float)A[1]; f.oz = (float)A[2]; f.l1 = (float)l1;
    f.ux = (float)u[0]; f.uy = (float)u[1]; f.uz = (float)u[2]; f.x2 = (float)x2;
    f.wx = (float)w[0]; f.wy = (float)w[1]; f.wz = (float)w[2]; f.y2 = (float)y2;
    return f;
}

namespace {

// One child of a 4-wide node: its box (FP32, already rounded outward) and
// its reference in the binary forest.
struct WideChild { float lo[3], hi[3]; int32_t ref; };

// The 4-wide forest from the binary FP32 forest: wide node of binary node b
// holds b's children, each internal child replaced by its two children.
// Numbered breadth-first over all domains (roots first), like the binary
// forest, so the top levels form a prefix.  Leaf references keep their code.
void build_wide(DeviceLayouts &L, int dim) {
    auto children = [&](int32_t b, WideChild out[2]) {
        if (dim == 3) {
            const NodeF3 &n = L.nf3[(size_t)b];
            out[0] = WideChild{{n.l_lox, n.l_loy, n.l_loz}, {n.l_hix, n.l_hiy, n.l_hiz}, n.left};
            out[1] = WideChild{{n.r_lox, n.r_loy, n.r_loz}, {n.r_hix, n.r_hiy, n.r_hiz}, n.right};
        } else {
            const NodeF2 &n = L.nf2[(size_t)b];
            out[0] = WideChild{{n.l_lox, n.l_loy, 0.0f}, {n.l_hix, n.l_hiy, 0.0f}, n.left};
            out[1] = WideChild{{n.r_lox, n.r_loy, 0.0f}, {n.r_hix, n.r_hiy, 0.0f}, n.right};
        }
    };
    std::vector<int32_t> queue, level;  // binary internal nodes of the wide nodes
    std::vector<std::vector<WideChild>> kids;
    for (DomainInfo &D : L.domains) {
        if (D.root_ref >= 0) {
            D.root_ref_w = (int32_t)queue.size();
            queue.push_back(D.root_ref);
            level.push_back(1);
        } else {
            D.root_ref_w = D.root_ref;  // a leaf root
        }
    }
    L.depth_w = 0;
    for (size_t h = 0; h < queue.size(); ++h) {
        WideChild c2[2], g2[2];
        children(queue[h], c2);
        std::vector<WideChild> k;
        for (const WideChild &c : c2) {
            if (c.ref < 0) {
                k.push_back(c);
                continue;
            }
            children(c.ref, g2);
            for (WideChild g : g2) {
                if (g.ref >= 0) {
                    const int32_t wi = (int32_t)queue.size();
                    queue.push_back(g.ref);
                    level.push_back(level[h] + 1);
                    g.ref = wi;
                }
                k.push_back(g);
            }
        }
        L.depth_w = std::max(L.depth_w, level[h]);
        kids.push_back(std::move(k));
    }
    const float inf = std::numeric_limits<float>::infinity();
    if (dim == 3) L.nw3.assign(queue.size(), NodeF3W{}); else L.nw2.assign(queue.size(), NodeF2W{});
    for (size_t h = 0; h < queue.size(); ++h) {
        for (int j = 0; j < 4; ++j) {
            const bool used = j < (int)kids[h].size();
            const WideChild c = used ? kids[h][(size_t)j]
                                     : WideChild{{inf, inf, inf}, {-inf, -inf, -inf}, 0};
            if (dim == 3) {
                NodeF3W &n = L.nw3[h];
                n.lox[j] = c.lo[0]; n.loy[j] = c.lo[1]; n.loz[j] = c.lo[2];
                n.hix[j] = c.hi[0]; n.hiy[j] = c.hi[1]; n.hiz[j] = c.hi[2];
                n.ref[j] = c.ref;
            } else {
                NodeF2W &n = L.nw2[h];
                n.lox[j] = c.lo[0]; n.loy[j] = c.lo[1];
                n.hix[j] = c.hi[0]; n.hiy[j] = c.hi[1];
                n.ref[j] = c.ref;
            }
        }
    }
}

}  // namespace

DeviceLayouts make_layouts(const std::vector<const HostMesh *> &meshes,
                           const std::vector<const HostBVH *> &bvhs) {
    DeviceLayouts L;
    const int dim = meshes[0]->dim;
    const size_t nd = meshes.size();
    // global offsets
    L.domains.resize(nd);
    int64_t vo = 0, po = 0;
    int32_t no = 0;
    size_t n_nodes = 0;
    for (size_t d = 0; d < nd; ++d) {
        DomainInfo &D = L.domains[d];
        D.vert_off = vo; D.n_verts = meshes[d]->n_verts();
        D.prim_off = po; D.n_prims = meshes[d]->n_prims();
        D.node_off = no;
        D.depth = bvhs[d]->depth;
        L.depth = std::max(L.depth, D.depth);
        vo += D.n_verts; po += D.n_prims;
        no += (int32_t)bvhs[d]->nodes.size();
        n_nodes += bvhs[d]->nodes.size();
        // bounding box as the reference computes it (src/mesh.cpp:296-322)
        const auto &V = meshes[d]->verts;
        double lo[3] = {V[0], V[1], dim == 3 ? V[2] : 0.0};
        double hi[3] = {lo[0], lo[1], lo[2]};
        for (int64_t i = 1; i < D.n_verts; ++i)
            for (int k = 0; k < dim; ++k) {
                const double v = V[(size_t)i * dim + k];
                if (v < lo[k]) lo[k] = v;
                else if (v > hi[k]) hi[k] = v;
            }
        for (int k = 0; k < 3; ++k) { D.bounds[2 * k] = lo[k]; D.bounds[2 * k + 1] = hi[k]; }
        for (double v : V) D.coord_scale = std::max(D.coord_scale, std::fabs(v));
        L.coord_scale = std::max(L.coord_scale, D.coord_scale);
    }
    const size_t ns = (size_t)po;
    L.slot_prim.resize(ns);
    if (dim == 2) {
        L.nd2.resize(n_nodes); L.segd.resize(ns); L.segf.resize(ns);
    } else {
        L.nd3.resize(n_nodes); L.trid.resize(ns); L.trif.resize(ns);
    }
    for (size_t d = 0; d < nd; ++d) {
        const HostMesh &m = *meshes[d];
        const HostBVH &b = *bvhs[d];
        const DomainInfo &D = L.domains[d];
        auto V = [&](int32_t vi, int k) { return m.verts[(size_t)vi * dim + k]; };
        for (size_t i = 0; i < b.nodes.size(); ++i) {
            const RefNode &r = b.nodes[i];
            // leaf: first global slot; internal: global right child
            const int32_t idx = r.count > 0 ? (int32_t)(r.idx + D.prim_off) : r.idx + D.node_off;
            if (dim == 2) L.nd2[D.node_off + i] = NodeD2{r.lo[0], r.lo[1], r.hi[0], r.hi[1], r.count, idx};
            else L.nd3[D.node_off + i] = NodeD3{r.lo[0], r.lo[1], r.lo[2], r.hi[0], r.hi[1], r.hi[2], r.count, idx};
        }
        for (size_t s = 0; s < b.slots.size(); ++s) {
            const size_t gs = (size_t)D.prim_off + s;
            const int32_t p = b.slots[s];
            L.slot_prim[gs] = (int32_t)(p + D.prim_off);
            if (dim == 2) {
                const int32_t a = m.prims[2 * p], c = m.prims[2 * p + 1];
                L.segd[gs] = SegD{V(a, 0), V(a, 1), V(c, 0), V(c, 1)};
                const float ax = (float)V(a, 0), ay = (float)V(a, 1);
                const float vx = (float)V(c, 0) - ax, vy = (float)V(c, 1) - ay;
                L.segf[gs] = SegF{ax, ay, vx, vy};
            } else {
                const int32_t i0 = m.prims[3 * p], i1 = m.prims[3 * p + 1], i2 = m.prims[3 * p + 2];
                TriD t{};
                t.v0x = V(i0, 0); t.v0y = V(i0, 1); t.v0z = V(i0, 2);
                t.e1x = V(i1, 0) - t.v0x; t.e1y = V(i1, 1) - t.v0y; t.e1z = V(i1, 2) - t.v0z;
                t.e2x = V(i2, 0) - t.v0x; t.e2y = V(i2, 1) - t.v0y; t.e2z = V(i2, 2) - t.v0z;
                // cache fill of src/bvh.cpp:199-203
                t.a = t.e1x * t.e1x + t.e1y * t.e1y + t.e1z * t.e1z;
                t.b = t.e1x * t.e2x + t.e1y * t.e2y + t.e1z * t.e2z;
                t.c = t.e2x * t.e2x + t.e2y * t.e2y + t.e2z * t.e2z;
                t.det = t.a * t.c - t.b * t.b;
                t.denom = t.a - 2 * t.b + t.c;
                L.trid[gs] = t;
                L.trif[gs] = triangle_frame(m, i0, i1, i2);
            }
        }
    }

    // FP32 child-box layout: internal nodes of the whole forest in
    // breadth-first order (all roots first), global leaf slots.
    std::vector<int32_t> remap(n_nodes, -1);
    int32_t ni = 0;
    {
        std::vector<int32_t> queue;
        queue.reserve(n_nodes);
        for (size_t d = 0; d < nd; ++d) {
            const int32_t r = L.domains[d].node_off;
            if ((dim == 2 ? L.nd2[r].count : L.nd3[r].count) == 0) queue.push_back(r);
        }
        for (size_t h = 0; h < queue.size(); ++h) {
            const int32_t i = queue[h];
            remap[i] = ni++;
            const int32_t l = i + 1, r = dim == 2 ? L.nd2[i].idx : L.nd3[i].idx;
            if ((dim == 2 ? L.nd2[l].count : L.nd3[l].count) == 0) queue.push_back(l);
            if ((dim == 2 ? L.nd2[r].count : L.nd3[r].count) == 0) queue.push_back(r);
        }
    }
    auto count_of = [&](int32_t i) { return dim == 2 ? L.nd2[i].count : L.nd3[i].count; };
    auto idx_of = [&](int32_t i) { return dim == 2 ? L.nd2[i].idx : L.nd3[i].idx; };
    auto ref_of = [&](int32_t node) -> int32_t {
        if (count_of(node) > 0) return -(1 + ((idx_of(node) << 3) | count_of(node)));
        return remap[node];
    };
    for (size_t d = 0; d < nd; ++d) L.domains[d].root_ref = ref_of(L.domains[d].node_off);
    if (dim == 2) L.nf2.resize(ni); else L.nf3.resize(ni);
    for (size_t i = 0; i < n_nodes; ++i) {
        if (count_of((int32_t)i) > 0) continue;
        const int32_t l = (int32_t)i + 1, rr = idx_of((int32_t)i);
        if (dim == 2) {
            const NodeD2 &A = L.nd2[l], &B = L.nd2[rr];
            NodeF2 n{};
            n.l_lox = round_down(A.lox); n.l_loy = round_down(A.loy);
            n.l_hix = round_up(A.hix); n.l_hiy = round_up(A.hiy);
            n.r_lox = round_down(B.lox); n.r_loy = round_down(B.loy);
            n.r_hix = round_up(B.hix); n.r_hiy = round_up(B.hiy);
            n.left = ref_of(l);
            n.right = ref_of(rr);
            L.nf2[remap[i]] = n;
        } else {
            const NodeD3 &A = L.nd3[l], &B = L.nd3[rr];
            NodeF3 n{};
            n.l_lox = round_down(A.lox); n.l_loy = round_down(A.loy); n.l_loz = round_down(A.loz);
            n.l_hix = round_up(A.hix); n.l_hiy = round_up(A.hiy); n.l_hiz = round_up(A.hiz);
            n.r_lox = round_down(B.lox); n.r_loy = round_down(B.loy); n.r_loz = round_down(B.loz);
            n.r_hix = round_up(B.hix); n.r_hiy = round_up(B.hiy); n.r_hiz = round_up(B.hiz);
            n.left = ref_of(l);
            n.right = ref_of(rr);
            L.nf3[remap[i]] = n;
        }
    }
    build_wide(L, dim);
    return L;
}

}  // namespace wosx
