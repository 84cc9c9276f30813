// bvh_build.cpp -- host BVH construction (reference split rule) and the
// flattening into the FP64 reference-order and FP32 child-box layouts.
#include "bvh_build.hpp"

#include <algorithm>
#include <cmath>

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

DeviceLayouts make_layouts(const HostMesh &m, const HostBVH &b) {
    DeviceLayouts L;
    const int dim = m.dim;
    const size_t nn = b.nodes.size(), ns = b.slots.size();
    double scale = 1.0;
    for (double v : m.verts) scale = std::max(scale, std::fabs(v));
    L.coord_scale = scale;
    auto V = [&](int32_t vi, int k) { return m.verts[(size_t)vi * dim + k]; };

    if (dim == 2) {
        L.nd2.resize(nn);
        for (size_t i = 0; i < nn; ++i) {
            const RefNode &r = b.nodes[i];
            L.nd2[i] = NodeD2{r.lo[0], r.lo[1], r.hi[0], r.hi[1], r.count, r.idx};
        }
        L.segd.resize(ns);
        L.segf.resize(ns);
        L.seg_inv_vv.resize(ns);
        for (size_t s = 0; s < ns; ++s) {
            const int32_t p = b.slots[s];
            const int32_t a = m.prims[2 * p], c = m.prims[2 * p + 1];
            L.segd[s] = SegD{V(a, 0), V(a, 1), V(c, 0), V(c, 1)};
            const float ax = (float)V(a, 0), ay = (float)V(a, 1);
            const float vx = (float)V(c, 0) - ax, vy = (float)V(c, 1) - ay;
            L.segf[s] = SegF{ax, ay, vx, vy};
            const float vv = vx * vx + vy * vy;
            L.seg_inv_vv[s] = vv > 0.0f ? 1.0f / vv : 0.0f;
        }
    } else {
        L.nd3.resize(nn);
        for (size_t i = 0; i < nn; ++i) {
            const RefNode &r = b.nodes[i];
            L.nd3[i] = NodeD3{r.lo[0], r.lo[1], r.lo[2], r.hi[0], r.hi[1], r.hi[2], r.count, r.idx};
        }
        L.trid.resize(ns);
        L.trif.resize(ns);
        for (size_t s = 0; s < ns; ++s) {
            const int32_t p = b.slots[s];
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
            L.trid[s] = t;
            TriF f{};
            f.v0x = (float)t.v0x; f.v0y = (float)t.v0y; f.v0z = (float)t.v0z;
            f.e1x = (float)t.e1x; f.e1y = (float)t.e1y; f.e1z = (float)t.e1z;
            f.e2x = (float)t.e2x; f.e2y = (float)t.e2y; f.e2z = (float)t.e2z;
            f.a = (float)t.a; f.b = (float)t.b; f.c = (float)t.c;
            f.det = (float)t.det; f.denom = (float)t.denom;
            f.inv_det = t.det != 0.0 ? (float)(1.0 / t.det) : 0.0f;
            L.trif[s] = f;
        }
    }

    // FP32 child-box layout: the internal nodes of the reference tree in
    // breadth-first order, so the top levels are one contiguous prefix that a
    // block can stage in LDS.
    auto ref_of = [&](int32_t node, const std::vector<int32_t> &remap) -> int32_t {
        const RefNode &r = b.nodes[node];
        if (r.count > 0) return -(1 + ((r.idx << 3) | r.count));
        return remap[node];
    };
    std::vector<int32_t> remap(nn, -1);
    int32_t ni = 0;
    {
        std::vector<int32_t> queue;
        queue.reserve(nn);
        if (b.nodes[0].count == 0) queue.push_back(0);
        for (size_t h = 0; h < queue.size(); ++h) {
            const int32_t i = queue[h];
            remap[i] = ni++;
            const int32_t l = i + 1, r = b.nodes[i].idx;
            if (b.nodes[l].count == 0) queue.push_back(l);
            if (b.nodes[r].count == 0) queue.push_back(r);
        }
    }
    L.root_ref = ref_of(0, remap);
    if (dim == 2) L.nf2.resize(ni); else L.nf3.resize(ni);
    for (size_t i = 0; i < nn; ++i) {
        const RefNode &r = b.nodes[i];
        if (r.count > 0) continue;
        const int32_t l = (int32_t)i + 1, rr = r.idx;
        const RefNode &A = b.nodes[l], &B = b.nodes[rr];
        if (dim == 2) {
            NodeF2 n{};
            n.l_lox = round_down(A.lo[0]); n.l_loy = round_down(A.lo[1]);
            n.l_hix = round_up(A.hi[0]); n.l_hiy = round_up(A.hi[1]);
            n.r_lox = round_down(B.lo[0]); n.r_loy = round_down(B.lo[1]);
            n.r_hix = round_up(B.hi[0]); n.r_hiy = round_up(B.hi[1]);
            n.left = ref_of(l, remap);
            n.right = ref_of(rr, remap);
            L.nf2[remap[i]] = n;
        } else {
            NodeF3 n{};
            n.l_lox = round_down(A.lo[0]); n.l_loy = round_down(A.lo[1]); n.l_loz = round_down(A.lo[2]);
            n.l_hix = round_up(A.hi[0]); n.l_hiy = round_up(A.hi[1]); n.l_hiz = round_up(A.hi[2]);
            n.r_lox = round_down(B.lo[0]); n.r_loy = round_down(B.lo[1]); n.r_loz = round_down(B.lo[2]);
            n.r_hix = round_up(B.hi[0]); n.r_hiy = round_up(B.hi[1]); n.r_hiz = round_up(B.hi[2]);
            n.left = ref_of(l, remap);
            n.right = ref_of(rr, remap);
            L.nf3[remap[i]] = n;
        }
    }
    return L;
}

}  // namespace wosx
