// wos_device.hpp -- device functions of the walk-on-spheres kernels (gfx950).
//
// FP64 functions restate the reference operation for operation and are
// compiled with -ffp-contract=off, so the classifier's location codes and the
// FP64 walker's per-walk values match the reference / oracle arithmetic.
// FP32 functions are the redesigned fast path (child-box BVH, conservative
// radii); statistics stay FP64.
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>

#include "../../include/wos_hip.h"
#include "bvh_build.hpp"
#include "wos_common.hpp"

namespace wosx {

// --------------------------------------------------------------------------
// distance grid (FP32 walker): a uniform grid of cubic cells per domain.
// Cell record: the distance dc from the cell centre to the boundary as FP32
// bits with the low 7 mantissa bits cleared (so still a lower bound) and
// replaced by the length (0..127) of the cell's primitive list, which starts
// at glist[first].  The list holds every primitive within dc + 2 hd of the
// centre (hd = half cell diagonal, padded), which contains the nearest
// primitive of every point within hd of the centre; entries carry their
// distance from the centre (rounded down) and are sorted by it, so a scan
// from p can stop at the first entry with dmin - |p - c| >= its best.  A
// cell whose list would exceed the build's cap stores none (length 0), and
// so does a cell whose bound dc already answers every relaxed query.  Entry
// 0 (always tested: the scan starts from an infinite best) and the bound of
// entry 1 are repeated in the cell record, so a scan that entry 1 cannot
// improve never reads the list.
// --------------------------------------------------------------------------
struct GridInfo {
    float ox, oy, oz;     // corner of cell (0, 0, 0)
    float h, inv_h;       // cell size and its reciprocal
    float hd;             // padded half diagonal used for the lists
    int32_t nx, ny, nz;   // cells per axis (nz = 1 in 2D)
    int32_t cell_off;     // first cell of this domain in the cell array
};
struct alignas(16) GridCellF {
    uint32_t dcbits;   // FP32 bits of dc, low 7 bits = list length
    uint32_t first;    // first list entry; bit 31 (kGridPlanar): the list's
                       // triangles lie in one plane, kept in MeshView::planes
    int32_t slot0;     // leaf slot of list entry 0 (nearest to the centre), -1 if none
    float dmin1;       // list entry 1's distance from the centre, +inf if none
};
struct GridEntry {
    int32_t slot;      // leaf slot of the primitive
    float dmin;        // its distance from the cell centre, rounded down
};
constexpr int kGridListMax = 127;
constexpr uint32_t kGridLenMask = 127u;
constexpr uint32_t kGridPlanar = 0x80000000u;

// --------------------------------------------------------------------------
// kernel-side views
// --------------------------------------------------------------------------
// per-domain constants of the FP32 walker
struct alignas(8) DomainF {
    int32_t root32;       // FP32 root reference
    float shrink_abs;     // absolute margin of FP32 radii
};

struct MeshView {
    int dim;
    int n_prims;
    int n_nodes;
    int depth;
    // FP64, reference order
    const NodeD2 *nd2;
    const NodeD3 *nd3;
    const SegD *segd;
    const TriD *trid;
    const int32_t *slot_prim;  // leaf slot -> primitive id
    const int32_t *prim_bid;   // primitive -> boundary id (2D)
    const int32_t *prims;      // primitive -> vertex ids (dim per primitive)
    const double *verts;       // vertex coordinates (dim per vertex)
    // FP32, child-box layout (binary and 4-wide)
    const NodeF2 *nf2;
    const NodeF3 *nf3;
    const NodeF2W *nw2;
    const NodeF3W *nw3;
    const SegF *segf;
    const TriF *trif;
    // distance grid (nullptr when off)
    const GridInfo *grids;     // per domain
    const GridCellF *cells;
    const GridEntry *glist;    // cell lists
    // per cell (3D grids): for a planar cell the plane (unit normal n, offset
    // d = n . x on the plane) of its list's triangles
    const float4 *planes;
    const DomainF *doms;       // per domain
};

struct EqView {
    int kind, source_mode, source_kind, boundary_kind;
    double alpha, src_value, cx, cy, cz, src_radius;
    double bnd[WOS_MAX_BND_VALUES];
    int n_bnd;
    const double *vertex_values;
    // GREEN_MIS source proposal: Green-draw probability, per-axis width,
    // normalisation (integral of the unnormalised Gaussian over the box or
    // the plane) and the truncation box (used when mis_box != 0)
    double mis_beta, mis_sigma, mis_norm;
    double box[4];
    int mis_box;
    // WOS_SRC_FIELD: the previous field (device copy, or null), its grid and
    // interpolation, and the heat terms (wos_field_source of the header)
    const double *field;
    int fnx, fny, finterp, fclamp;
    double fb[4];
    double hist_scale, heat_scale, heat_div;
    int n_heat;
    wos_heat_term heat[WOS_MAX_HEAT_TERMS];
};

// start of a walk, produced by the classifier for every point
struct StartInfo {
    double x, y, z;       // the point
    double r;             // distance to the boundary
    double nx, ny, nz;    // nearest boundary point
    double g;             // g at the nearest point (Boundary value, zero-step walks)
    double s, t;          // nearest primitive parameters
    int32_t prim, bid;
    int32_t slot;         // leaf slot of the nearest primitive
    int32_t root64;       // the point's domain: FP64 root node
    int32_t root32;       // the point's domain: FP32 root reference
    float shrink_abs;     // the point's domain: absolute margin of FP32 radii
    uint32_t key0, key1;  // walk-stream key
    int32_t domain;       // the point's domain index
    int32_t pad;
    uint32_t skey0, skey1;  // 2D source-stream seed of the reference streams
    uint32_t pad1, pad2;
};

// the FP32 walker's copy of a start (two 16 B loads), one per inside point
struct alignas(16) StartF {
    float x, y, z;        // the point, rounded to FP32
    float r;              // distance to the boundary, rounded to FP32
    int32_t slot;         // leaf slot of the nearest primitive
    int32_t domain;
    uint32_t key0, key1;  // walk-stream key
};

// nearest-point query result (FP64)
struct NpqD {
    double d2;
    double px, py, pz;
    double s, t;
    int32_t prim;
    int32_t slot;
};

// --------------------------------------------------------------------------
// FP64 primitives (src/npq.cpp:8-29, :52-189)
// --------------------------------------------------------------------------
__device__ __forceinline__ double npq_seg_d(const SegD &g, double px, double py, double &nx,
                                            double &ny, double &par) {
    const double vx = g.s1x - g.s0x, vy = g.s1y - g.s0y;
    const double wx = px - g.s0x, wy = py - g.s0y;
    const double vv = vx * vx + vy * vy;
    const double wv = wx * vx + wy * vy;
    par = (vv > 0.0) ? wv / vv : 0.0;
    if (par < 0.0) par = 0.0;
    if (par > 1.0) par = 1.0;
    nx = g.s0x + par * vx;
    ny = g.s0y + par * vy;
    const double dx = px - nx, dy = py - ny;
    return dx * dx + dy * dy;
}

__device__ __forceinline__ double eberly_d(const TriD &T, double px, double py, double pz,
                                           double &nx, double &ny, double &nz, double &so,
                                           double &to) {
    const double a = T.a, b = T.b, c = T.c, det = T.det, denom = T.denom;
    const double bpx = T.v0x - px, bpy = T.v0y - py, bpz = T.v0z - pz;
    const double d = T.e1x * bpx + T.e1y * bpy + T.e1z * bpz;
    const double e = T.e2x * bpx + T.e2y * bpy + T.e2z * bpz;
    double s = b * e - c * d;
    double t = b * d - a * e;
    if (s + t <= det) {
        if (s < 0) {
            if (t < 0) {
                if (d < 0) {
                    t = 0;
                    if (-d >= a) s = 1; else s = -d / a;
                } else {
                    s = 0;
                    if (e >= 0) t = 0; else if (-e >= c) t = 1; else t = -e / c;
                }
            } else {
                s = 0;
                if (e >= 0) t = 0; else if (-e >= c) t = 1; else t = -e / c;
            }
        } else if (t < 0) {
            t = 0;
            if (d >= 0) s = 0; else if (-d >= a) s = 1; else s = -d / a;
        } else {
            s /= det;
            t /= det;
        }
    } else {
        if (s < 0) {
            const double bd = b + d, ce = c + e;
            if (ce > bd) {
                const double numer = ce - bd;
                if (numer >= denom) s = 1; else s = numer / denom;
                t = 1 - s;
            } else {
                s = 0;
                if (ce <= 0) t = 1; else if (e >= 0) t = 0; else t = -e / c;
            }
        } else if (t < 0) {
            const double be = b + e, ad = a + d;
            if (ad > be) {
                const double numer = ad - be;
                if (numer >= denom) t = 1; else t = numer / denom;
                s = 1 - t;
            } else {
                t = 0;
                if (ad <= 0) s = 1; else if (d >= 0) s = 0; else s = -d / a;
            }
        } else {
            const double numer = (c + e) - (b + d);
            if (numer <= 0) s = 0;
            else if (numer >= denom) s = 1;
            else s = numer / denom;
            t = 1 - s;
        }
    }
    nx = T.v0x + s * T.e1x + t * T.e2x;
    ny = T.v0y + s * T.e1y + t * T.e2y;
    nz = T.v0z + s * T.e1z + t * T.e2z;
    const double dx = nx - px, dy = ny - py, dz = nz - pz;
    so = s;
    to = t;
    return dx * dx + dy * dy + dz * dz;
}

// src/bvh.cpp:111-121 (dmax of include/wos/fastmath.hpp:11)
__device__ __forceinline__ double dmaxd(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double box_d2_2(const NodeD2 &n, double px, double py) {
    const double dx = dmaxd(0.0, dmaxd(n.lox - px, px - n.hix));
    const double dy = dmaxd(0.0, dmaxd(n.loy - py, py - n.hiy));
    return dx * dx + dy * dy;
}
__device__ __forceinline__ double box_d2_3(const NodeD3 &n, double px, double py, double pz) {
    const double dx = dmaxd(0.0, dmaxd(n.lox - px, px - n.hix));
    const double dy = dmaxd(0.0, dmaxd(n.loy - py, py - n.hiy));
    const double dz = dmaxd(0.0, dmaxd(n.loz - pz, pz - n.hiz));
    return dx * dx + dy * dy + dz * dz;
}

// Closest point, reference traversal.  `stk` is this lane's LDS stack
// (stride = block size).  The reference's search (src/bvh.cpp:295-387): pop
// a node, skip it when
// its box is no nearer than the best so far, test a leaf's primitives in
// order, or push the far child and then the near one (near = smaller box
// distance, ties to the left).  Here the near child is not pushed and
// popped again: the search continues into it with the record it has just
// loaded and the box distance it has just formed (nothing changes the best
// distance in between, so the skip test is the same), which saves one
// dependent load per internal node; the far child waits on the stack as in
// the reference.  Same nodes, same order, same result (2D).
__device__ __forceinline__ NpqD npq_ref_2d(const MeshView &m, int root, double px, double py,
                                           int *stk, int stride) {
    NpqD r;
    r.d2 = INFINITY; r.px = r.py = r.pz = 0.0; r.s = r.t = 0.0; r.prim = -1; r.slot = -1;
    int sp = 0;
    NodeD2 n = m.nd2[root];
    int curr = root;
    double nd = box_d2_2(n, px, py);
    for (;;) {
        if (nd < r.d2) {
            if (n.count > 0) {
                for (int i = 0; i < n.count; ++i) {
                    const int slot = n.idx + i;
                    double nx, ny, par;
                    const double d = npq_seg_d(m.segd[slot], px, py, nx, ny, par);
                    if (d < r.d2) {
                        r.d2 = d; r.px = nx; r.py = ny; r.s = par; r.prim = m.slot_prim[slot]; r.slot = slot;
                    }
                }
            } else {
                const int l = curr + 1, rr = n.idx;
                const NodeD2 nl = m.nd2[l], nr = m.nd2[rr];
                const double ld = box_d2_2(nl, px, py);
                const double rd = box_d2_2(nr, px, py);
                if (ld <= rd) { stk[(sp++) * stride] = rr; curr = l; n = nl; nd = ld; }
                else { stk[(sp++) * stride] = l; curr = rr; n = nr; nd = rd; }
                continue;
            }
        }
        if (sp == 0) break;
        curr = stk[(--sp) * stride];
        n = m.nd2[curr];
        nd = box_d2_2(n, px, py);
    }
    return r;
}

// In 3D the reference's loop as it stands: there the FP64 walker, which
// spills registers, measured the continuation 8 % slower on the 200k-triangle
// plate (5 % faster on the icosphere).  budget > 0: the search gives up after
// that many nodes and returns slot -2 (the caller searches again otherwise).
__device__ __forceinline__ NpqD npq_ref_3d(const MeshView &m, int root, double px, double py,
                                           double pz, int *stk, int stride, int budget = 0) {
    NpqD r;
    r.d2 = INFINITY; r.px = r.py = r.pz = 0.0; r.s = r.t = 0.0; r.prim = -1; r.slot = -1;
    int sp = 1;
    stk[0] = root;
    int pops = 0;
    while (sp > 0) {
        if (budget > 0 && ++pops > budget) {
            r.slot = -2;
            return r;
        }
        const int curr = stk[(--sp) * stride];
        const NodeD3 n = m.nd3[curr];
        if (box_d2_3(n, px, py, pz) >= r.d2) continue;
        if (n.count > 0) {
            for (int i = 0; i < n.count; ++i) {
                const int slot = n.idx + i;
                double nx, ny, nz, s, t;
                const double d = eberly_d(m.trid[slot], px, py, pz, nx, ny, nz, s, t);
                if (d < r.d2) {
                    r.d2 = d; r.px = nx; r.py = ny; r.pz = nz; r.s = s; r.t = t;
                    r.prim = m.slot_prim[slot];
                    r.slot = slot;
                }
            }
        } else {
            const int l = curr + 1, rr = n.idx;
            const double ld = box_d2_3(m.nd3[l], px, py, pz);
            const double rd = box_d2_3(m.nd3[rr], px, py, pz);
            if (ld <= rd) { stk[(sp++) * stride] = rr; stk[(sp++) * stride] = l; }
            else { stk[(sp++) * stride] = l; stk[(sp++) * stride] = rr; }
        }
    }
    return r;
}

// Resumable form of the reference's search (npq_ref_2d / npq_ref_3d): the
// same pops, box tests against the best so far, pushes (far child first,
// near child popped next) and primitive tests in slot order with the strict
// update, one unit per call: a node popped and processed, or one primitive
// of the leaf being scanned.  A lane's sequence of decisions is the
// reference's, so its result is the reference's bit for bit, while the
// lanes of a wavefront whose searches differ in length advance together.
struct QueryD {
    NpqD r;      // best so far
    int sp;      // entries on the lane's LDS stack
    int li, lc;  // leaf being scanned: next slot, primitives left
};

__device__ __forceinline__ void query_d_begin(QueryD &q, int root, int *stk) {
    q.r.d2 = INFINITY; q.r.px = q.r.py = q.r.pz = 0.0; q.r.s = q.r.t = 0.0;
    q.r.prim = -1; q.r.slot = -1;
    q.sp = 1;
    stk[0] = root;
    q.li = 0;
    q.lc = 0;
}

// One unit; returns whether the search goes on.
template <int DIM>
__device__ __forceinline__ bool query_d_step(QueryD &q, const MeshView &m, double px, double py,
                                             double pz, int *stk, int stride) {
    if (q.lc > 0) {
        const int slot = q.li;
        if constexpr (DIM == 3) {
            double nx, ny, nz, s, t;
            const double d = eberly_d(m.trid[slot], px, py, pz, nx, ny, nz, s, t);
            if (d < q.r.d2) {
                q.r.d2 = d; q.r.px = nx; q.r.py = ny; q.r.pz = nz; q.r.s = s; q.r.t = t;
                q.r.prim = m.slot_prim[slot];
                q.r.slot = slot;
            }
        } else {
            double nx, ny, par;
            const double d = npq_seg_d(m.segd[slot], px, py, nx, ny, par);
            if (d < q.r.d2) {
                q.r.d2 = d; q.r.px = nx; q.r.py = ny; q.r.s = par;
                q.r.prim = m.slot_prim[slot];
                q.r.slot = slot;
            }
        }
        ++q.li;
        --q.lc;
        return q.lc > 0 || q.sp > 0;
    }
    const int curr = stk[(--q.sp) * stride];
    int count, idx;
    double bd;
    if constexpr (DIM == 3) {
        const NodeD3 n = m.nd3[curr];
        bd = box_d2_3(n, px, py, pz);
        count = n.count; idx = n.idx;
    } else {
        const NodeD2 n = m.nd2[curr];
        bd = box_d2_2(n, px, py);
        count = n.count; idx = n.idx;
    }
    if (bd < q.r.d2) {
        if (count > 0) {
            q.li = idx;
            q.lc = count;
        } else {
            const int l = curr + 1, rr = idx;
            double ld, rd;
            if constexpr (DIM == 3) {
                ld = box_d2_3(m.nd3[l], px, py, pz);
                rd = box_d2_3(m.nd3[rr], px, py, pz);
            } else {
                ld = box_d2_2(m.nd2[l], px, py);
                rd = box_d2_2(m.nd2[rr], px, py);
            }
            if (ld <= rd) { stk[(q.sp++) * stride] = rr; stk[(q.sp++) * stride] = l; }
            else { stk[(q.sp++) * stride] = l; stk[(q.sp++) * stride] = rr; }
        }
    }
    return q.lc > 0 || q.sp > 0;
}

// --------------------------------------------------------------------------
// FP64 classification (src/inside.cpp)
// --------------------------------------------------------------------------
// src/inside.cpp:117-137 over every segment of the domain (order independent)
__device__ __forceinline__ int winding_number_d(const MeshView &m, int s0, int s1, double px,
                                                double py) {
    int wn = 0;
    for (int s = s0; s < s1; ++s) {
        const SegD g = m.segd[s];
        const double area = (g.s1x - g.s0x) * (py - g.s0y) - (px - g.s0x) * (g.s1y - g.s0y);
        if (g.s0y <= py) {
            if (g.s1y > py && area > 0) wn++;
        } else {
            if (g.s1y <= py && area < 0) wn--;
        }
    }
    return wn;
}

#define WOS_INTERSECTION_EPSILON 1e-9  // src/inside.cpp:14

// src/inside.cpp:140-182 on the cached (v0, edge1, edge2): 0 miss, 1 hit, 2 ambiguous
__device__ __forceinline__ int moller_trumbore_d(const TriD &T, double px, double py, double pz,
                                                 double rx, double ry, double rz) {
    const double e1x = T.e1x, e1y = T.e1y, e1z = T.e1z;
    const double e2x = T.e2x, e2y = T.e2y, e2z = T.e2z;
    const double sx = px - T.v0x, sy = py - T.v0y, sz = pz - T.v0z;
    const double qx = ry * e2z - rz * e2y, qy = rz * e2x - rx * e2z, qz = rx * e2y - ry * e2x;
    const double det = e1x * qx + e1y * qy + e1z * qz;
    const double nx = e1y * e2z - e1z * e2y, ny = e1z * e2x - e1x * e2z, nz = e1x * e2y - e1y * e2x;
    const double normal_length = sqrt(nx * nx + ny * ny + nz * nz);
    const double det_tolerance = WOS_INTERSECTION_EPSILON * normal_length;
    if (fabs(det) <= det_tolerance) {
        if (normal_length == 0.0) return 0;
        double edge_scale = 1.0;
        const double l1 = sqrt(e1x * e1x + e1y * e1y + e1z * e1z);
        const double l2 = sqrt(e2x * e2x + e2y * e2y + e2z * e2z);
        if (l1 > edge_scale) edge_scale = l1;
        if (l2 > edge_scale) edge_scale = l2;
        const double plane_distance = fabs(nx * sx + ny * sy + nz * sz) / normal_length;
        return plane_distance <= WOS_INTERSECTION_EPSILON * edge_scale ? 2 : 0;
    }
    const double u = (sx * qx + sy * qy + sz * qz) / det;
    if (u < -WOS_INTERSECTION_EPSILON || u - 1.0 > WOS_INTERSECTION_EPSILON) return 0;
    const double cx = sy * e1z - sz * e1y, cy = sz * e1x - sx * e1z, cz = sx * e1y - sy * e1x;
    const double v = (rx * cx + ry * cy + rz * cz) / det;
    if (v < -WOS_INTERSECTION_EPSILON || u + v - 1.0 > WOS_INTERSECTION_EPSILON) return 0;
    const double t = (e2x * cx + e2y * cy + e2z * cz) / det;
    if (t <= WOS_INTERSECTION_EPSILON) return 0;
    const double w = 1.0 - u - v;
    if (u <= WOS_INTERSECTION_EPSILON || v <= WOS_INTERSECTION_EPSILON ||
        w <= WOS_INTERSECTION_EPSILON)
        return 2;
    return 1;
}

// src/inside.cpp:27-51
__device__ __forceinline__ bool update_ray_interval(double origin, double direction, double box_min,
                                                    double box_max, double &t_min, double &t_max) {
    const double ext = fabs(box_max - box_min);
    const double extent_scale = ext > 1.0 ? ext : 1.0;
    double coord = 1.0;
    if (fabs(origin) > coord) coord = fabs(origin);
    if (fabs(box_min) > coord) coord = fabs(box_min);
    if (fabs(box_max) > coord) coord = fabs(box_max);
    const double padding = WOS_INTERSECTION_EPSILON * extent_scale + 16.0 * DBL_EPSILON * coord;
    box_min -= padding;
    box_max += padding;
    if (direction == 0.0) return origin >= box_min && origin <= box_max;
    double t0 = (box_min - origin) / direction;
    double t1 = (box_max - origin) / direction;
    if (t0 > t1) { const double tmp = t0; t0 = t1; t1 = tmp; }
    if (t0 > t_min) t_min = t0;
    if (t1 < t_max) t_max = t1;
    return t_min <= t_max;
}

__device__ __forceinline__ bool ray_hits_box(const NodeD3 &n, double px, double py, double pz,
                                             double rx, double ry, double rz) {
    double tmin = 0.0, tmax = INFINITY;
    return update_ray_interval(px, rx, n.lox, n.hix, tmin, tmax) &&
           update_ray_interval(py, ry, n.loy, n.hiy, tmin, tmax) &&
           update_ray_interval(pz, rz, n.loz, n.hiz, tmin, tmax);
}

// src/inside.cpp:78-111; returns hit count, sets ambiguous
__device__ __forceinline__ int ray_cast_d(const MeshView &m, int root, double px, double py,
                                          double pz, double rx, double ry, double rz,
                                          bool &ambiguous, int *stk, int stride) {
    int hits = 0, sp = 1;
    stk[0] = root;
    ambiguous = false;
    while (sp > 0) {
        const int ni = stk[(--sp) * stride];
        const NodeD3 n = m.nd3[ni];
        if (!ray_hits_box(n, px, py, pz, rx, ry, rz)) continue;
        if (n.count > 0) {
            for (int i = 0; i < n.count; ++i) {
                const int h = moller_trumbore_d(m.trid[n.idx + i], px, py, pz, rx, ry, rz);
                if (h == 2) { ambiguous = true; return hits; }
                if (h == 1) ++hits;
            }
        } else {
            stk[(sp++) * stride] = n.idx;
            stk[(sp++) * stride] = ni + 1;
        }
    }
    return hits;
}

// --------------------------------------------------------------------------
// special functions (FP64: the oracle's methods; FP32: same, fewer terms)
// --------------------------------------------------------------------------
// Elementary functions of the walk-step arithmetic (weights, Green
// functions, source draws) by precision.  FP64, the reference arithmetic,
// calls the C library and divides.  FP32 uses the hardware exp2, log2,
// reciprocal, sin and cos instructions with a few FMAs; relative errors stay
// near 1e-7 (exp of an argument x: about |x| 6e-8 more), far below the
// statistical tolerance of the FP32 walker.
template <typename T> struct Fx;
template <> struct Fx<double> {
    static __device__ __forceinline__ double exp(double x) { return ::exp(x); }
    static __device__ __forceinline__ double log(double x) { return ::log(x); }
    static __device__ __forceinline__ double expm1(double x) { return ::expm1(x); }
    static __device__ __forceinline__ double cbrt(double x) { return ::cbrt(x); }
    static __device__ __forceinline__ double cos(double x) { return ::cos(x); }
    static __device__ __forceinline__ double sin(double x) { return ::sin(x); }
    static __device__ __forceinline__ double div(double a, double b) { return a / b; }
    static __device__ __forceinline__ double sqrt(double x) { return ::sqrt(x); }
};
template <> struct Fx<float> {
    static __device__ __forceinline__ float exp(float x) {
        return __builtin_amdgcn_exp2f(x * 1.44269504088896341f);
    }
    static __device__ __forceinline__ float log(float x) {
        return __builtin_amdgcn_logf(x) * 0.693147180559945309f;
    }
    // e^x - 1: for |x| < 1/4 the Taylor polynomial of degree 7 (remainder
    // below x^8/8! < 2e-9 |x|), else e^x - 1 (cancellation at most 1/(1 - e^-1/4))
    static __device__ __forceinline__ float expm1(float x) {
        float p = __builtin_fmaf(x, 1.0f / 5040.0f, 1.0f / 720.0f);
        p = __builtin_fmaf(x, p, 1.0f / 120.0f);
        p = __builtin_fmaf(x, p, 1.0f / 24.0f);
        p = __builtin_fmaf(x, p, 1.0f / 6.0f);
        p = __builtin_fmaf(x, p, 0.5f);
        p = __builtin_fmaf(x, p, 1.0f);
        p *= x;
        return fabsf(x) < 0.25f ? p : exp(x) - 1.0f;
    }
    // cube root of x > 0: y = 2^(log2(x) / 3), then one Newton step
    // y - (y^3 - x) / (3 y^2); 0 for x <= 0
    static __device__ __forceinline__ float cbrt(float x) {
        const float y = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(x) * (1.0f / 3.0f));
        const float y2 = y * y;
        const float r = __builtin_fmaf(-(y2 * y - x), __builtin_amdgcn_rcpf(3.0f * y2), y);
        return x > 0.0f ? r : 0.0f;
    }
    static __device__ __forceinline__ float cos(float x) { return __cosf(x); }
    static __device__ __forceinline__ float sin(float x) { return __sinf(x); }
    static __device__ __forceinline__ float div(float a, float b) {
        return a * __builtin_amdgcn_rcpf(b);
    }
    // the square root instruction (within 1 ulp, not correctly rounded)
    static __device__ __forceinline__ float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
};

template <typename T> struct Lim;
template <> struct Lim<double> { static constexpr double eps = DBL_EPSILON; };
template <> struct Lim<float> { static constexpr float eps = FLT_EPSILON; };

template <typename T>
__device__ __forceinline__ T asymptotic_sum(T x, bool alternating) {
    T sum = 1, term = 1, previous = (T)INFINITY;
    for (int k = 1; k <= 24; ++k) {
        const T odd = (T)(2 * k - 1);
        T factor = odd * odd / ((T)8 * x * (T)k);
        if (alternating) factor = -factor;
        term *= factor;
        const T mag = fabs(term);
        if (mag > previous) break;
        sum += term;
        if (mag <= Lim<T>::eps * fabs(sum)) break;
        previous = mag;
    }
    return sum;
}

// e^-x I0(x): power series below 50 (A&S 9.6.12), Hankel series above
// (include/wos/bessel.hpp:40-48)
template <typename T>
__device__ __forceinline__ T i0_scaled(T x) {
    x = fabs(x);
    if (x < (T)50) {
        const T q = (T)0.25 * x * x;
        T term = 1, sum = 1;
        for (int k = 1; k < 500; ++k) {
            term *= q / ((T)k * (T)k);
            sum += term;
            if (term <= (T)1e-17 * sum) break;
        }
        return exp(-x) * sum;
    }
    return asymptotic_sum<T>(x, false) / sqrt((T)2 * (T)kPi * x);
}

// e^x K0(x): A&S 9.6.13 series for x <= 2, trapezoidal rule on
// int_0^inf exp(-2x sinh^2(t/2)) dt (A&S 9.6.24) below 50, Hankel above
template <typename T>
__device__ __forceinline__ T k0_scaled(T x) {
    if (x < (T)0) return (T)NAN;
    if (x == (T)0) return (T)INFINITY;
    if (x <= (T)2) {
        const T q = (T)0.25 * x * x;
        T term = 1, hk = 0, sum = 0, i0 = 1;
        for (int k = 1; k < 200; ++k) {
            term *= q / ((T)k * (T)k);
            hk += (T)1 / (T)k;
            i0 += term;
            sum += term * hk;
            if (term <= (T)1e-18) break;
        }
        const T k0 = -(log((T)0.5 * x) + (T)0.57721566490153286060651209) * i0 + sum;
        return exp(x) * k0;
    }
    if (x < (T)50) {
        // step below 0.5/sqrt(x): the integrand is about exp(-x t^2 / 2)
        const T h = fmin((T)0.125, (T)0.5 / sqrt(x));
        T sum = (T)0.5;
        for (int i = 1; i < 4000; ++i) {
            const T sh = sinh((T)0.5 * h * (T)i);
            const T v = exp((T)-2 * x * sh * sh);
            sum += v;
            if (v < (T)1e-18) break;
        }
        return h * sum;
    }
    return sqrt((T)kPi / ((T)2 * x)) * asymptotic_sum<T>(x, true);
}

// FP32 I0 and K0 by the polynomial approximations of Abramowitz & Stegun
// 9.8.1-9.8.2 and 9.8.5-9.8.6 (relative error below 2e-7, three FP32 ulps):
// a few FMAs in place of the series and quadrature loops the FP64 path
// shares with the reference.
__device__ __forceinline__ float i0_small_f(float x) {     // I0(x), |x| <= 3.75 (9.8.1)
    const float t = x * (1.0f / 3.75f), t2 = t * t;
    float p = 0.0045813f;
    p = __builtin_fmaf(p, t2, 0.0360768f);
    p = __builtin_fmaf(p, t2, 0.2659732f);
    p = __builtin_fmaf(p, t2, 1.2067492f);
    p = __builtin_fmaf(p, t2, 3.0899424f);
    p = __builtin_fmaf(p, t2, 3.5156229f);
    return __builtin_fmaf(p, t2, 1.0f);
}
__device__ __forceinline__ float log_i0_f(float x) {
    x = fabsf(x);
    if (x <= 3.75f) return Fx<float>::log(i0_small_f(x));
    // sqrt(x) e^-x I0(x) = P(3.75 / x) (9.8.2)
    const float u = 3.75f * __builtin_amdgcn_rcpf(x);
    float p = 0.00392377f;
    p = __builtin_fmaf(p, u, -0.01647633f);
    p = __builtin_fmaf(p, u, 0.02635537f);
    p = __builtin_fmaf(p, u, -0.02057706f);
    p = __builtin_fmaf(p, u, 0.00916281f);
    p = __builtin_fmaf(p, u, -0.00157565f);
    p = __builtin_fmaf(p, u, 0.00225319f);
    p = __builtin_fmaf(p, u, 0.01328592f);
    p = __builtin_fmaf(p, u, 0.39894228f);
    return x + Fx<float>::log(p) - 0.5f * Fx<float>::log(x);
}
__device__ __forceinline__ float log_k0_f(float x) {
    if (x < 0.0f) return NAN;
    if (x == 0.0f) return INFINITY;
    if (x <= 2.0f) {
        // K0(x) = -ln(x/2) I0(x) + P((x/2)^2) (9.8.5)
        const float h = 0.5f * x, h2 = h * h;
        float p = 0.00000740f;
        p = __builtin_fmaf(p, h2, 0.00010750f);
        p = __builtin_fmaf(p, h2, 0.00262698f);
        p = __builtin_fmaf(p, h2, 0.03488590f);
        p = __builtin_fmaf(p, h2, 0.23069756f);
        p = __builtin_fmaf(p, h2, 0.42278420f);
        p = __builtin_fmaf(p, h2, -0.57721566f);
        return Fx<float>::log(__builtin_fmaf(-Fx<float>::log(h), i0_small_f(x), p));
    }
    // sqrt(x) e^x K0(x) = P(2 / x) (9.8.6)
    const float u = 2.0f * __builtin_amdgcn_rcpf(x);
    float p = 0.00053208f;
    p = __builtin_fmaf(p, u, -0.00251540f);
    p = __builtin_fmaf(p, u, 0.00587872f);
    p = __builtin_fmaf(p, u, -0.01062446f);
    p = __builtin_fmaf(p, u, 0.02189568f);
    p = __builtin_fmaf(p, u, -0.07832358f);
    p = __builtin_fmaf(p, u, 1.25331414f);
    return -x + Fx<float>::log(p) - 0.5f * Fx<float>::log(x);
}

template <typename T> __device__ __forceinline__ T log_i0(T x) {
    if constexpr (sizeof(T) == 4) return log_i0_f(x);
    x = fabs(x);
    return x + log(i0_scaled<T>(x));
}
template <typename T> __device__ __forceinline__ T log_k0(T x) {
    if constexpr (sizeof(T) == 4) return log_k0_f(x);
    if (x < (T)0) return (T)NAN;
    if (x == (T)0) return (T)INFINITY;
    return -x + log(k0_scaled<T>(x));
}

// include/wos/poisson.hpp:12-38
template <typename T> __device__ __forceinline__ T poisson_green_2d(T R, T r) {
    if (!(r > (T)0) || r > R || R <= (T)0) return (T)NAN;
    if (r == R) return (T)0;
    return Fx<T>::div(Fx<T>::log(Fx<T>::div(R, r)), (T)2 * (T)kPi);
}
template <typename T> __device__ __forceinline__ T poisson_green_3d(T R, T r) {
    if (!(r > (T)0) || r > R || R <= (T)0) return (T)NAN;
    if (r == R) return (T)0;
    return Fx<T>::div(R - r, (T)4 * (T)kPi * r * R);
}
// include/wos/screened.hpp:14-91
template <typename T> __device__ __forceinline__ T screened_weight_2d(T alpha, T R) {
    const T x = alpha * R;
    if (x == (T)0) return (T)1;
    if (x < (T)1e-3) {
        const T x2 = x * x;
        return (T)1 - x2 / (T)4 + (T)3 * x2 * x2 / (T)64;
    }
    return exp(-log_i0<T>(x));
}
template <typename T> __device__ __forceinline__ T screened_green_2d(T alpha, T R, T r) {
    if (!(r > (T)0) || r > R || R <= (T)0) return (T)NAN;
    if (r == R) return (T)0;
    const T x = alpha * r, X = alpha * R;
    if (X < (T)1e-5) return poisson_green_2d<T>(R, r);
    const T log_a = log_k0<T>(x);
    const T log_b = log_i0<T>(x) + log_k0<T>(X) - log_i0<T>(X);
    T delta = log_b - log_a;
    if (delta > (T)0) delta = (T)0;
    return exp(log_a) * (-expm1(delta)) / ((T)2 * (T)kPi);
}
template <typename T> __device__ __forceinline__ T screened_mass_2d(T alpha, T R) {
    const T x = alpha * R;
    if (x < (T)1e-3) {
        const T x2 = x * x;
        return R * R * ((T)1 / (T)4 - (T)3 * x2 / (T)64 + (T)19 * x2 * x2 / (T)2304);
    }
    return -expm1(-log_i0<T>(x)) / (alpha * alpha);
}
template <typename T> __device__ __forceinline__ T screened_weight_3d(T alpha, T R) {
    const T x = alpha * R;
    if (x == (T)0) return (T)1;
    if (x < (T)1e-3) {
        const T x2 = x * x;
        return (T)1 - Fx<T>::div(x2, (T)6) + Fx<T>::div((T)7 * x2 * x2, (T)360);
    }
    return Fx<T>::div((T)2 * x * Fx<T>::exp(-x), -Fx<T>::expm1((T)-2 * x));
}
template <typename T> __device__ __forceinline__ T screened_green_3d(T alpha, T R, T r) {
    if (!(r > (T)0) || r > R || R <= (T)0) return (T)NAN;
    if (r == R) return (T)0;
    const T x = alpha * r, X = alpha * R;
    if (X < (T)1e-5) return poisson_green_3d<T>(R, r);
    const T ratio = Fx<T>::div(Fx<T>::exp(-x) * (-Fx<T>::expm1((T)-2 * (X - x))),
                               -Fx<T>::expm1((T)-2 * X));
    return Fx<T>::div(ratio, (T)4 * (T)kPi * r);
}
// The 3D screened kernel's terms of one walk step at X = alpha R (FP32
// walker): e^-X and q = 1 - e^-2X, shared by the step's weight X / sinh X =
// 2 X e^-X / q and its source term's Green function, whose ratio
// sinh(alpha (R - r)) / sinh X = e^-(alpha r) (1 - e^-2 alpha (R - r)) / q.
// The same formulas as screened_weight_3d / screened_green_3d.
struct Scr3 { float X, eX, q; };
__device__ __forceinline__ Scr3 scr3_terms(float alpha, float R) {
    const float X = alpha * R;
    return Scr3{X, Fx<float>::exp(-X), -Fx<float>::expm1(-2.0f * X)};
}
__device__ __forceinline__ float scr3_weight(const Scr3 &S) {
    if (S.X < 1e-3f) {
        const float x2 = S.X * S.X;
        return __builtin_fmaf(x2, __builtin_fmaf(x2, 7.0f / 360.0f, -1.0f / 6.0f), 1.0f);
    }
    return 2.0f * S.X * S.eX * __builtin_amdgcn_rcpf(S.q);
}
// G at distance 0 < r <= R from the centre (the walker's source samples)
__device__ __forceinline__ float scr3_green(const Scr3 &S, float alpha, float R, float r) {
    if (!(r < R)) return 0.0f;
    if (S.X < 1e-5f) return poisson_green_3d<float>(R, r);
    const float x = alpha * r;
    const float ratio = Fx<float>::exp(-x) * (-Fx<float>::expm1(-2.0f * (S.X - x))) *
                        __builtin_amdgcn_rcpf(S.q);
    return ratio * __builtin_amdgcn_rcpf(4.0f * (float)kPi * r);
}

// include/wos/fastmath.hpp:15-29: the series of J0, stopped at the first
// term below 1e-8 (the reference's own truncation)
template <typename T> __device__ __forceinline__ T bessel_j0_series(T x) {
    const T q = (T)0.25 * x * x;
    T term = 1, sum = 1;
    for (int k = 1; k < 200; ++k) {
        term *= Fx<T>::div(q, (T)(k * k));
        if (term < (T)1e-8) return sum;
        sum += (k % 2 == 0 ? term : -term);
    }
    return sum;
}
// Helmholtz weights (src/equations/helmholtz.cpp:36-41, :70-75)
template <typename T> __device__ __forceinline__ T helmholtz_weight_2d(T k, T R) {
    const T prod = R * k;
    if (prod < (T)1e-8) return (T)1;
    return Fx<T>::div((T)1, bessel_j0_series<T>(prod));
}
template <typename T> __device__ __forceinline__ T helmholtz_weight_3d(T k, T R) {
    const T prod = R * k;
    if (prod < (T)1e-8) return (T)1;
    return Fx<T>::div(prod, Fx<T>::sin(prod));
}
template <typename T> __device__ __forceinline__ T screened_mass_3d(T alpha, T R) {
    const T x = alpha * R;
    if (x < (T)1e-3) {
        const T x2 = x * x;
        return R * R * ((T)1 / (T)6 - Fx<T>::div((T)7 * x2, (T)360) +
                        Fx<T>::div((T)31 * x2 * x2, (T)15120));
    }
    return Fx<T>::div((T)1 - screened_weight_3d<T>(alpha, R), alpha * alpha);
}

// the walk weight of one step: screened Poisson or Helmholtz (alpha = k)
template <int DIM, typename T> __device__ __forceinline__ T step_weight(int kind, T alpha, T R) {
    if (kind == WOS_EQ_HELMHOLTZ)
        return DIM == 2 ? helmholtz_weight_2d<T>(alpha, R) : helmholtz_weight_3d<T>(alpha, R);
    return DIM == 2 ? screened_weight_2d<T>(alpha, R) : screened_weight_3d<T>(alpha, R);
}

// --------------------------------------------------------------------------
// alpha-derivatives of the 3D screened weight and Green function, as the
// derivatives of the functions above (series branch included), in FP64
// --------------------------------------------------------------------------
// d/dalpha log w3(alpha R)
__device__ __forceinline__ double dlog_w3_dalpha(double alpha, double R) {
    const double x = alpha * R;
    if (x == 0.0) return 0.0;
    if (x < 1e-3) {
        const double x2 = x * x;
        return R * (-x / 3.0 + 7.0 * x2 * x / 90.0) / (1.0 - x2 / 6.0 + 7.0 * x2 * x2 / 360.0);
    }
    return R * (1.0 / x - 1.0 / tanh(x));
}
// 2 y / expm1(2 alpha y), with its limit 1/alpha at y = 0
__device__ __forceinline__ double h_screen(double alpha, double y) {
    const double z = 2.0 * alpha * y;
    return z == 0.0 ? 1.0 / alpha : 2.0 * y / expm1(z);
}
// d/dalpha log G3(alpha; R, r)
__device__ __forceinline__ double dlog_g3_dalpha(double alpha, double R, double r) {
    if (alpha * R < 1e-5) return 0.0;
    return -r + h_screen(alpha, R - r) - h_screen(alpha, R);
}

// The FP32 replay forms the same two derivatives from
// phi(y) = 2y / expm1(2 alpha y) - 1/alpha + y, the screened kernel's
// deviation from its linear part: d/dalpha log w3 = -phi(R) and
// d/dalpha log G3 = phi(R - r) - phi(R), which cancels no 1/alpha terms
// (the forms above subtract quantities of size 1/alpha).  With
// z = 2 alpha y, alpha phi = z/expm1(z) - 1 + z/2 = z^2/12 - z^4/720 +
// z^6/30240 - z^8/1209600 + ... (Bernoulli series, used for |z| < 1).
__device__ __forceinline__ float phi_screen_f(float alpha, float y) {
    const float z = 2.0f * alpha * y;
    const float z2 = z * z;
    float s = __builtin_fmaf(z2, -1.0f / 1209600.0f, 1.0f / 30240.0f);
    s = __builtin_fmaf(z2, s, -1.0f / 720.0f);
    s = __builtin_fmaf(z2, s, 1.0f / 12.0f);
    s *= z2;
    const float d = Fx<float>::div(z, Fx<float>::expm1(z)) - 1.0f + 0.5f * z;
    return Fx<float>::div(fabsf(z) < 1.0f ? s : d, alpha);
}
__device__ __forceinline__ float dlog_w3_dalpha_f(float alpha, float R) {
    return alpha * R == 0.0f ? 0.0f : -phi_screen_f(alpha, R);
}
__device__ __forceinline__ float dlog_g3_dalpha_f(float alpha, float R, float r) {
    if (alpha * R < 1e-5f) return 0.0f;
    return phi_screen_f(alpha, R - r) - phi_screen_f(alpha, R);
}

// Per-walk derivatives accumulated by the adjoint replay.
struct WalkDer {
    double d_alpha;           // d value / d alpha
    double d_q;               // d value / d src_value
    double d_sigma;           // d value / d src_radius
    double d_c[3];            // d value / d src_center
    double L;                 // running d log(weight) / d alpha
};

// source-term derivatives for a term S = base * f(y) (base = weight x
// volume x G or weight x mass); dlogG = d log G / d alpha (0 if unused).
// T is the replay's precision: the Gaussian factor exp(-|y - c|^2 / s^2) of
// d/dQ is formed in it.
template <typename T>
__device__ __forceinline__ void add_source_der(WalkDer &D, const EqView &e, int dim, double base,
                                               double S, double yx, double yy, double yz,
                                               double dlogG) {
    if (e.source_kind == WOS_SRC_GAUSSIAN) {
        const double dx = yx - e.cx, dy = yy - e.cy, dz = dim == 3 ? yz - e.cz : 0.0;
        const double s2 = e.src_radius * e.src_radius;
        const double d2 = dx * dx + dy * dy + dz * dz;
        D.d_q += base * (double)Fx<T>::exp(-Fx<T>::div((T)d2, (T)s2));
        D.d_sigma += S * 2.0 * d2 / (s2 * e.src_radius);
        D.d_c[0] += S * 2.0 * dx / s2;
        D.d_c[1] += S * 2.0 * dy / s2;
        D.d_c[2] += S * 2.0 * dz / s2;
    } else if (e.src_value != 0.0) {
        D.d_q += S / e.src_value;
    }
    D.d_alpha += S * (D.L + dlogG);
}

// --------------------------------------------------------------------------
// WOS_SRC_FIELD: interpolation of the previous field and the heat terms
// --------------------------------------------------------------------------
template <typename T> __device__ __forceinline__ T clampt(T v, T lo, T hi) {
    return v < lo ? lo : (hi < v ? hi : v);
}

// FieldSampler2D::sample (apps/welding/include/welding/field_sampler.hpp:51-117):
// grid coordinates clamped to the grid, then bilinear on the cell or tensor
// cubic Lagrange on the 4 x 4 nodes around it (shifted inside at the edges)
template <typename T>
__device__ T field_sample(const EqView &e, T x, T y) {
    const int nx = e.fnx, ny = e.fny;
    const T gx = clampt<T>((x - (T)e.fb[0]) / ((T)e.fb[1] - (T)e.fb[0]) * (T)(nx - 1), (T)0,
                           (T)(nx - 1));
    const T gy = clampt<T>((y - (T)e.fb[2]) / ((T)e.fb[3] - (T)e.fb[2]) * (T)(ny - 1), (T)0,
                           (T)(ny - 1));
    const double *v = e.field;
    if (e.finterp == WOS_INTERP_CUBIC && nx >= 4 && ny >= 4) {
        const int ix = min(max((int)floor(gx) - 1, 0), nx - 4);
        const int iy = min(max((int)floor(gy) - 1, 0), ny - 4);
        T wx[4], wy[4];
        const T cxs[2] = {gx - (T)ix, gy - (T)iy};
        for (int k = 0; k < 2; ++k) {
            const T c = cxs[k];
            T *w = k == 0 ? wx : wy;
            w[0] = -(c - (T)1) * (c - (T)2) * (c - (T)3) / (T)6;
            w[1] = c * (c - (T)2) * (c - (T)3) / (T)2;
            w[2] = -c * (c - (T)1) * (c - (T)3) / (T)2;
            w[3] = c * (c - (T)1) * (c - (T)2) / (T)6;
        }
        T value = (T)0;
        for (int di = 0; di < 4; ++di) {
            const double *r = v + (size_t)(ix + di) * ny + iy;
            T row = (T)0;
            for (int dj = 0; dj < 4; ++dj) row += wy[dj] * (T)r[dj];
            value += wx[di] * row;
        }
        return value;
    }
    const int i0 = (int)floor(gx), j0 = (int)floor(gy);
    const int i1 = min(i0 + 1, nx - 1), j1 = min(j0 + 1, ny - 1);
    const T tx = gx - (T)i0, ty = gy - (T)j0;
    const T v00 = (T)v[(size_t)i0 * ny + j0], v10 = (T)v[(size_t)i1 * ny + j0];
    const T v01 = (T)v[(size_t)i0 * ny + j1], v11 = (T)v[(size_t)i1 * ny + j1];
    const T lower = ((T)1 - tx) * v00 + tx * v10;
    const T upper = ((T)1 - tx) * v01 + tx * v11;
    return ((T)1 - ty) * lower + ty * upper;
}

// f = history + heat (level2_transient_heat.cpp:255-264,
// level3_moving_heat.cpp:189-194, level0_cooling.cpp:194-196); each heat
// term a exp(-(dx^2 + dy^2) / 2) / b as GaussianHeatSource2D (:38-41, :109-113)
// and MovingGaussianStepSource2D (:75-88) evaluate it
template <typename T>
__device__ T field_source_value(const EqView &e, T x, T y) {
    T history = (T)0;
    if (e.field) {
        const T v = field_sample<T>(e, x, y);
        history = (T)e.hist_scale * (e.fclamp ? (v < (T)0 ? (T)0 : v) : v);
    }
    T heat = (T)0;
    if (e.heat_scale != 0.0 && e.n_heat > 0) {
        T sum = (T)0;
        for (int c = 0; c < e.n_heat; ++c) {
            const wos_heat_term &h = e.heat[c];
            const T dx = (x - (T)h.cx) / (T)h.sx, dy = (y - (T)h.cy) / (T)h.sy;
            sum += (T)h.a * Fx<T>::exp((T)-0.5 * (dx * dx + dy * dy)) / (T)h.b;
        }
        heat = (T)e.heat_scale * sum / (T)e.heat_div;
    }
    return history + heat;
}

// --------------------------------------------------------------------------
// equation functors (src/equations/*.cpp) driven by the descriptor
// --------------------------------------------------------------------------
// FIELD = true compiles the WOS_SRC_FIELD case (kernels launched for it only)
template <typename T, bool FIELD = false>
__device__ __forceinline__ T eq_source(const EqView &e, T x, T y, T z, int dim) {
    const T dx = x - (T)e.cx, dy = y - (T)e.cy;
    T acc = dx * dx + dy * dy;
    if (dim == 3) { const T dz = z - (T)e.cz; acc = acc + dz * dz; }
    switch (e.source_kind) {
    case WOS_SRC_DISK_SQ: return acc <= (T)e.src_radius * (T)e.src_radius ? (T)e.src_value : (T)0;
    case WOS_SRC_DISK_SQRT: return sqrt(acc) <= (T)e.src_radius ? (T)e.src_value : (T)0;
    case WOS_SRC_CONSTANT: return (T)e.src_value;
    case WOS_SRC_GAUSSIAN:
        return (T)e.src_value * Fx<T>::exp(-Fx<T>::div(acc, (T)e.src_radius * (T)e.src_radius));
    case WOS_SRC_FIELD:
        if constexpr (FIELD) return field_source_value<T>(e, x, y);
        else return (T)0;
    default: return (T)0;
    }
}

template <typename T>
__device__ __forceinline__ bool eq_source_may_intersect(const EqView &e, T cx, T cy, T R) {
    if (e.source_kind != WOS_SRC_DISK_SQ && e.source_kind != WOS_SRC_DISK_SQRT) return true;
    const T dx = cx - (T)e.cx, dy = cy - (T)e.cy;
    const T reach = R + (T)e.src_radius;
    return dx * dx + dy * dy <= reach * reach;
}

// g at a nearest point (FP64 always)
__device__ __forceinline__ double eq_boundary(const EqView &e, const MeshView &m, double x,
                                              double y, double z, int prim, double s, double t) {
    switch (e.boundary_kind) {
    case WOS_BND_X2Y2_ID0: {
        const int bid = prim >= 0 ? m.prim_bid[prim] : -1;
        return bid == 0 ? x * x - y * y : 0.0;
    }
    case WOS_BND_Y5: {
        const double r_sq = x * x + y * y;
        const double y50 = z * (8 * z * z * z * z - 40 * z * z * r_sq + 15 * r_sq * r_sq);
        const double y5m3 = (8 * z * z - r_sq) * y * (3 * x * x - y * y);
        return y50 + 4.0 * y5m3;
    }
    case WOS_BND_ID_CONST: {
        const int bid = prim >= 0 ? m.prim_bid[prim] : -1;
        return (bid >= 0 && bid < e.n_bnd) ? e.bnd[bid] : 0.0;
    }
    case WOS_BND_VERTEX: {
        if (prim < 0) return 0.0;
        const int32_t *pr = m.prims + (size_t)prim * m.dim;
        const double *vv = e.vertex_values;
        if (m.dim == 2) return (1.0 - s) * vv[pr[0]] + s * vv[pr[1]];
        return (1.0 - s - t) * vv[pr[0]] + s * vv[pr[1]] + t * vv[pr[2]];
    }
    case WOS_BND_SIN3X_COS2Y: return sin(3.0 * x) * cos(2.0 * y);
    case WOS_BND_ONE: return 1.0;
    default: return 0.0;
    }
}

// --------------------------------------------------------------------------
// FP32 child-box traversal
// --------------------------------------------------------------------------
__device__ __forceinline__ float fmaxf3(float a, float b, float c) { return fmaxf(fmaxf(a, b), c); }
__device__ __forceinline__ float clamp01(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }

// squared distance point -> box, FP32
__device__ __forceinline__ float box_d2f(float lox, float loy, float loz, float hix, float hiy,
                                         float hiz, float px, float py, float pz) {
    const float dx = fmaxf3(0.0f, lox - px, px - hix);
    const float dy = fmaxf3(0.0f, loy - py, py - hiy);
    const float dz = fmaxf3(0.0f, loz - pz, pz - hiz);
    return __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
}
__device__ __forceinline__ float box_d2f2(float lox, float loy, float hix, float hiy, float px,
                                          float py) {
    const float dx = fmaxf3(0.0f, lox - px, px - hix);
    const float dy = fmaxf3(0.0f, loy - py, py - hiy);
    return __builtin_fmaf(dx, dx, dy * dy);
}

// Squared distance point -> triangle in FP32 without branches.  The nearest
// point is the plane projection when its barycentrics lie in the triangle,
// otherwise the nearest of the three clamped edge projections; the minimum of
// the in-triangle projection (when valid) and the three edge candidates is
// that distance in every case, degenerate triangles included (their
// reciprocals are 0, which collapses a candidate onto a vertex).  The same
// minimum as the region walk of src/npq.cpp:52-189.
// Squared distance from p to a triangle in its frame record: p - o in
// frame coordinates (x, y, h), then h^2 plus the squared in-plane distance to
// the 2D triangle (0, 0), (l1, 0), (x2, y2): zero strictly inside it, else
// the nearest of its three edges (parameters clamped to [0, 1]).
__device__ __forceinline__ float tri_d2f(const TriF &T, float px, float py, float pz) {
    const float rx = px - T.ox, ry = py - T.oy, rz = pz - T.oz;
    const float x = __builtin_fmaf(rx, T.ux, __builtin_fmaf(ry, T.uy, rz * T.uz));
    const float y = __builtin_fmaf(rx, T.wx, __builtin_fmaf(ry, T.wy, rz * T.wz));
    const float nx = __builtin_fmaf(T.uy, T.wz, -T.uz * T.wy);
    const float ny = __builtin_fmaf(T.uz, T.wx, -T.ux * T.wz);
    const float nz = __builtin_fmaf(T.ux, T.wy, -T.uy * T.wx);
    const float h = __builtin_fmaf(rx, nx, __builtin_fmaf(ry, ny, rz * nz));
    const float yy = y * y;
    // edge (0, 0)-(l1, 0)
    const float xa = x - __builtin_amdgcn_fmed3f(x, 0.0f, T.l1);
    const float dA = __builtin_fmaf(xa, xa, yy);
    // edge (0, 0)-(x2, y2); the floor keeps a zero-length edge finite
    const float invC = __builtin_amdgcn_rcpf(fmaxf(__builtin_fmaf(T.x2, T.x2, T.y2 * T.y2), 1e-30f));
    const float tC = clamp01(__builtin_fmaf(x, T.x2, y * T.y2) * invC);
    const float xc = __builtin_fmaf(-tC, T.x2, x), yc = __builtin_fmaf(-tC, T.y2, y);
    const float dC = __builtin_fmaf(xc, xc, yc * yc);
    // edge (l1, 0)-(x2, y2)
    const float bx = T.x2 - T.l1, xl = x - T.l1;
    const float invB = __builtin_amdgcn_rcpf(fmaxf(__builtin_fmaf(bx, bx, T.y2 * T.y2), 1e-30f));
    const float tB = clamp01(__builtin_fmaf(xl, bx, y * T.y2) * invB);
    const float xb = __builtin_fmaf(-tB, bx, xl), yb = __builtin_fmaf(-tB, T.y2, y);
    const float dB = __builtin_fmaf(xb, xb, yb * yb);
    // strictly left of the three edges of the counter-clockwise triangle
    const float crossB = __builtin_fmaf(bx, y, -T.y2 * xl);
    const float crossC = __builtin_fmaf(T.y2, x, -T.x2 * y);
    const bool inside = fminf(y, fminf(crossB, crossC)) > 0.0f;
    const float inplane = inside ? 0.0f : fminf(dA, fminf(dB, dC));
    return __builtin_fmaf(h, h, inplane);
}

// (a zero-length segment gives 0 x inf = NaN, which the clamp turns into 0)
__device__ __forceinline__ float seg_d2f(const SegF &g, float px, float py) {
    const float wx = px - g.s0x, wy = py - g.s0y;
    const float inv_vv = __builtin_amdgcn_rcpf(__builtin_fmaf(g.vx, g.vx, g.vy * g.vy));
    const float par = clamp01(__builtin_fmaf(wx, g.vx, wy * g.vy) * inv_vv);
    const float dx = __builtin_fmaf(-par, g.vx, wx), dy = __builtin_fmaf(-par, g.vy, wy);
    return __builtin_fmaf(dx, dx, dy * dy);
}

template <int DIM>
__device__ __forceinline__ float leaf_d2f(const MeshView &m, int slot, float px, float py, float pz) {
    if constexpr (DIM == 3) return tri_d2f(m.trif[slot], px, py, pz);
    else return seg_d2f(m.segf[slot], px, py);
}

struct StackEntry { int32_t ref; float d2; };

// A lane's traversal stack: entries 0 .. klds-1 in LDS (`lds`, `lstride`
// entries apart), deeper ones in global memory (`glob`, `gstride` apart).
struct StackRef {
    StackEntry *lds;
    int lstride;
    StackEntry *glob;
    int gstride;
    int klds;
    __device__ __forceinline__ void put(int i, StackEntry e) const {
        if (i < klds) lds[i * lstride] = e;
        else glob[i * gstride] = e;
    }
    __device__ __forceinline__ StackEntry get(int i) const {
        return i < klds ? lds[i * lstride] : glob[i * gstride];
    }
};

// Resumable closest-point query of one lane.  A query advances by one unit
// per call of step(): one internal node (both child boxes) or one leaf
// primitive, so lanes of a wave whose queries differ in length by orders of
// magnitude stay busy together; the walker interleaves its walk steps with
// these units.  Order: near child first, far child with its box distance on
// the lane's LDS stack.
//
// Pruning: a box is opened while its squared distance is below cf x best.
// cf = 1 is the exact search.  cf < 1 skips boxes between cf x best and best;
// the smallest such box distance is kept in lb, so min(best, lb) is a lower
// bound of the true squared distance and best (reached by best_slot) an upper
// bound; both are equal whenever lb >= best.
struct QueryF {
    int32_t cur;      // internal node to open (when lcount == 0)
    int32_t lslot;    // next leaf primitive (its index in glist when indirect)
    int32_t lcount;   // leaf primitives left
    bool indirect;    // primitives come from a grid cell list
    float t;          // |p - cell centre| of a list scan
    int32_t sp;       // stack depth
    float best;       // smallest squared primitive distance so far
    float lb;         // smallest skipped box distance
    float cf;         // pruning factor
    int32_t best_slot;
    uint32_t ckey;    // 1 + index of the grid cell held in `cell`, 0 for none
    GridCellF cell;   // the last grid cell read (cells do not change during a launch)
    int32_t tslot;    // slot of the triangle in the lane's LDS copy (-1: none)
};

// A lane's LDS copy of one triangle record, three 16 B rows `stride`
// float4 apart (rows of consecutive lanes are adjacent: conflict-free).
struct TriCache {
    float4 *row;
    int stride;
};

// the triangle of `slot` from the lane's LDS copy, or from global memory
// (then copied there); the walker's steps near the boundary keep testing
// the same cell's first list entry
__device__ __forceinline__ TriF tri_cached(const MeshView &m, QueryF &q, const TriCache &tc,
                                           int slot) {
    float4 r0, r1, r2;
    if (slot == q.tslot) {
        r0 = tc.row[0]; r1 = tc.row[tc.stride]; r2 = tc.row[2 * tc.stride];
    } else {
        const float4 *g = reinterpret_cast<const float4 *>(m.trif + slot);
        r0 = g[0]; r1 = g[1]; r2 = g[2];
        tc.row[0] = r0; tc.row[tc.stride] = r1; tc.row[2 * tc.stride] = r2;
        q.tslot = slot;
    }
    TriF T;
    T.ox = r0.x; T.oy = r0.y; T.oz = r0.z; T.l1 = r0.w;
    T.ux = r1.x; T.uy = r1.y; T.uz = r1.z; T.x2 = r1.w;
    T.wx = r2.x; T.wy = r2.y; T.wz = r2.z; T.y2 = r2.w;
    return T;
}

// Set the next unit from a child reference.
__device__ __forceinline__ void query_enter(QueryF &q, int32_t ref) {
    if (ref >= 0) {
        q.cur = ref;
        q.lcount = 0;
    } else {
        const int32_t code = -ref - 1;
        q.lslot = code >> 3;
        q.lcount = code & 7;
    }
}

// What one query_step or query_begin did (bits of its return value).
constexpr uint32_t kQsMore = 1u;         // the query continues
constexpr uint32_t kQsNode = 2u;         // opened an internal node
constexpr uint32_t kQsLeaf = 4u;         // tested a BVH leaf primitive
constexpr uint32_t kQsEntryShift = 3u;   // cell-list entries read, 0..2 (bits 3-4)
constexpr uint32_t kQsPrimShift = 5u;    // cell-list primitives tested, 0..2 (bits 5-6)

// arr[index] of a small read-only per-domain table.  When every calling lane
// has the same index (always for a single mesh) the record is read through
// the constant address space with a wave-uniform index, i.e. by scalar
// loads, which keeps it off the vector memory path.
template <typename T>
__device__ __forceinline__ T domain_record(const T *arr, int index) {
    static_assert(sizeof(T) % 4 == 0, "whole 32-bit words");
    constexpr int kWords = (int)(sizeof(T) / 4);
    const int i0 = __builtin_amdgcn_readfirstlane(index);
    T rec;
    if (__ballot(index != i0) == 0) {
        typedef __attribute__((address_space(4))) const uint32_t ConstWord;
        const ConstWord *src = (const ConstWord *)(arr + i0);
        uint32_t w[kWords];
#pragma unroll
        for (int k = 0; k < kWords; ++k) w[k] = (uint32_t)__builtin_amdgcn_readfirstlane((int)src[k]);
        __builtin_memcpy(&rec, w, sizeof(rec));
    } else {
        rec = arr[index];
    }
    return rec;
}

// Begin a query at p in `domain`.  With a distance grid: with cf < 1 and
// PLANE, a planar cell (its list's triangles all in one plane) ends it at
// once with lb = (n . p - d)^2 and best = +inf: the nearest primitive of
// every point of a cell is on its list, and a triangle in the plane is no
// nearer than the plane, so |n . p - d| is a lower bound of the distance,
// and equal to it wherever p projects onto the face (a relaxed answer: a
// walk that it would end searches again exactly).  With cf < 1, a cell
// whose centre distance dc gives lb = dc - |p - c| and ub = dc' + |p - c|
// (dc' = dc rounded up) with lb^2 >= cf ub^2 ends it at once with
// best = ub^2, lb = lb^2; otherwise a cell that holds a primitive list turns
// it into a scan of that list (exact for cf = 1), whose entry 0 is tested
// here.  Otherwise the search starts at `root` with best from the primitive
// in q.best_slot (the previous step's nearest, -1 for none), which bounds it
// from the first node on.  Returns the kQs* bits; kQsMore is clear when the
// query is already complete.
template <int DIM, bool PLANE = false>
__device__ __forceinline__ uint32_t query_begin(QueryF &q, const MeshView &m, int32_t root,
                                                int domain, float cf, float px, float py,
                                                float pz, TriCache tc = TriCache{nullptr, 0}) {
    q.lb = INFINITY;
    q.cf = cf;
    q.sp = 0;
    q.indirect = false;
    if (m.grids) {
        const GridInfo G = domain_record(m.grids, domain);
        const float fx = (px - G.ox) * G.inv_h, fy = (py - G.oy) * G.inv_h;
        const float fz = DIM == 3 ? (pz - G.oz) * G.inv_h : 0.5f;
        const int ix = (int)floorf(fx), iy = (int)floorf(fy), iz = (int)floorf(fz);
        if (ix >= 0 && iy >= 0 && iz >= 0 && ix < G.nx && iy < G.ny && iz < G.nz) {
            // steps near the boundary are short and often stay in one cell
            const uint32_t key = (uint32_t)(G.cell_off + (ix * G.ny + iy) * G.nz + iz) + 1u;
            if (key != q.ckey) {
                q.cell = m.cells[key - 1u];
                q.ckey = key;
            }
            const GridCellF C = q.cell;
            if constexpr (PLANE && DIM == 3) {
                if (cf < 1.0f && (C.first & kGridPlanar) != 0u) {
                    const float4 P = m.planes[key - 1u];
                    const float h = fabsf(__builtin_fmaf(
                        P.x, px, __builtin_fmaf(P.y, py, __builtin_fmaf(P.z, pz, -P.w))));
                    q.lb = h * h;
                    q.best = INFINITY;
                    return 0u;
                }
            }
            const float dx = px - (G.ox + ((float)ix + 0.5f) * G.h);
            const float dy = py - (G.oy + ((float)iy + 0.5f) * G.h);
            const float dz = DIM == 3 ? pz - (G.oz + ((float)iz + 0.5f) * G.h) : 0.0f;
            const float t = Fx<float>::sqrt(__builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz)));
            if (cf < 1.0f) {
                // dc lost its low 7 mantissa bits: dc <= true <= dc (1 + 2^-15)
                const float dc = __uint_as_float(C.dcbits & ~kGridLenMask);
                const float lb = dc - t;
                const float ub = __builtin_fmaf(dc, 1.0f + 1.0f / 32768.0f, t);
                if (lb > 0.0f && lb * lb >= cf * (ub * ub)) {
                    q.best = ub * ub;
                    q.lb = lb * lb;
                    return 0u;
                }
            }
            const uint32_t len = C.dcbits & kGridLenMask;
            if (len != 0u) {
                // entry 0 now; the list only if entry 1's bound is below cf x best
                q.indirect = true;
                q.t = t;
                if constexpr (DIM == 3) {
                    if (tc.row) q.best = tri_d2f(tri_cached(m, q, tc, C.slot0), px, py, pz);
                    else q.best = leaf_d2f<DIM>(m, C.slot0, px, py, pz);
                } else {
                    q.best = leaf_d2f<DIM>(m, C.slot0, px, py, pz);
                }
                q.best_slot = C.slot0;
                q.lslot = (int32_t)(C.first & ~kGridPlanar) + 1;
                q.lcount = (int32_t)len - 1;
                if (q.lcount > 0) {
                    const float l1 = fmaxf(C.dmin1 - t, 0.0f);
                    const float l12 = l1 * l1;
                    if (l12 >= q.best * cf) {
                        if (l12 < q.best) q.lb = fminf(q.lb, l12);
                        q.lcount = 0;
                    }
                }
                const uint32_t did = 1u << kQsPrimShift;
                return q.lcount > 0 ? did | kQsMore : did;
            }
        }
    }
    q.best = q.best_slot >= 0 ? leaf_d2f<DIM>(m, q.best_slot, px, py, pz) : INFINITY;
    query_enter(q, root);
    return kQsMore;
}

// a leaf primitive's FP32 record, loaded apart from its distance test so
// that two loads can be in flight together
template <int DIM> struct LeafRec;
template <> struct LeafRec<3> { TriF t; };
template <> struct LeafRec<2> { SegF s; };
template <int DIM>
__device__ __forceinline__ LeafRec<DIM> leaf_load(const MeshView &m, int slot) {
    LeafRec<DIM> r;
    if constexpr (DIM == 3) r.t = m.trif[slot];
    else r.s = m.segf[slot];
    return r;
}
template <int DIM>
__device__ __forceinline__ float leaf_dist(const LeafRec<DIM> &r, float px, float py, float pz) {
    if constexpr (DIM == 3) return tri_d2f(r.t, px, py, pz);
    else return seg_d2f(r.s, px, py);
}

// One unit: an internal node, a BVH leaf primitive, or the next two entries
// of a cell list.  Returns the kQs* bits of what it did; kQsMore is clear
// when the query is complete.
template <int DIM, int W = 2>
__device__ __forceinline__ uint32_t query_step(QueryF &q, const MeshView &m, float px, float py,
                                               float pz, const StackRef &stk,
                                               const void *cache, int n_cached) {
    bool pop;
    uint32_t did = 0u;
    if (q.lcount > 0 && q.indirect) {
        // The next two sorted entries at once, both primitive records in
        // flight together.  An entry whose lower bound dmin - |p - c| reaches
        // cf x best ends the scan (all later ones are farther), its bound
        // kept in lb when below best (a relaxed stop).  The second entry is
        // tested when its bound is below the threshold the first was checked
        // against: a primitive the first one's result would have let the scan
        // skip cannot lower an exact minimum.
        // A scan starts at entry 1 of a list that begins at an odd index and
        // continues only by whole pairs, so lslot is even: one 16 B load
        // (past a list's end it reads padding or the next list, unused).
        const bool two = q.lcount > 1;
        const int4 pair = *reinterpret_cast<const int4 *>(m.glist + q.lslot);
        const GridEntry E0{pair.x, __int_as_float(pair.y)};
        const GridEntry E1{pair.z, __int_as_float(pair.w)};
        did |= (two ? 2u : 1u) << kQsEntryShift;
        const float thr = q.best * q.cf;
        const float l0 = fmaxf(E0.dmin - q.t, 0.0f), l1 = fmaxf(E1.dmin - q.t, 0.0f);
        const float l02 = l0 * l0, l12 = l1 * l1;
        if (l02 >= thr) {
            if (l02 < q.best) q.lb = fminf(q.lb, l02);
            q.lcount = 0;
        } else {
            // only the records that are tested are loaded, both in flight
            const bool t1 = two && l12 < thr;
            // (R1 is read only where t1 holds: no copy of R0 into it)
            const LeafRec<DIM> R0 = leaf_load<DIM>(m, E0.slot);
            LeafRec<DIM> R1;
            if (t1) R1 = leaf_load<DIM>(m, E1.slot);
            const float d0 = leaf_dist<DIM>(R0, px, py, pz);
            const float d1 = t1 ? leaf_dist<DIM>(R1, px, py, pz) : INFINITY;
            if (d0 < q.best) { q.best = d0; q.best_slot = E0.slot; }
            if (d1 < q.best) { q.best = d1; q.best_slot = E1.slot; }
            did |= (t1 ? 2u : 1u) << kQsPrimShift;
            if (two && !t1) {
                if (l12 < q.best) q.lb = fminf(q.lb, l12);
                q.lcount = 0;
            } else {
                const int used = t1 ? 2 : 1;
                q.lslot += used;
                q.lcount -= used;
            }
        }
        pop = q.lcount == 0;
    } else if (q.lcount > 0) {
        const int32_t slot = q.lslot;
        const float d = leaf_d2f<DIM>(m, slot, px, py, pz);
        if (d < q.best) { q.best = d; q.best_slot = slot; }
        ++q.lslot;
        --q.lcount;
        did |= kQsLeaf;
        pop = q.lcount == 0;
    } else if constexpr (W == 4) {
        // a 4-wide node: the children whose boxes are below cf x best, the
        // nearest entered and the others pushed farthest first
        float d[4];
        int32_t c[4];
        const int32_t ref = q.cur;
        if constexpr (DIM == 3) {
            const NodeF3W &n = ref < n_cached ? static_cast<const NodeF3W *>(cache)[ref] : m.nw3[ref];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                d[k] = box_d2f(n.lox[k], n.loy[k], n.loz[k], n.hix[k], n.hiy[k], n.hiz[k], px, py, pz);
                c[k] = n.ref[k];
            }
        } else {
            const NodeF2W &n = ref < n_cached ? static_cast<const NodeF2W *>(cache)[ref] : m.nw2[ref];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                d[k] = box_d2f2(n.lox[k], n.loy[k], n.hix[k], n.hiy[k], px, py);
                c[k] = n.ref[k];
            }
        }
        did |= kQsNode;
        const float thr = q.best * q.cf;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const bool hit = d[k] < thr;
            if (!hit && d[k] < q.best) q.lb = fminf(q.lb, d[k]);
            if (!hit) d[k] = INFINITY;
        }
        // sorting network on (distance, reference): ascending
        auto cx = [&](int i, int j) {
            const bool sw = d[j] < d[i];
            const float di = d[i], dj = d[j];
            const int32_t ci = c[i], cj = c[j];
            d[i] = sw ? dj : di; d[j] = sw ? di : dj;
            c[i] = sw ? cj : ci; c[j] = sw ? ci : cj;
        };
        cx(0, 1); cx(2, 3); cx(0, 2); cx(1, 3); cx(1, 2);
        pop = !(d[0] < INFINITY);
        if (!pop) {
            if (d[3] < INFINITY) stk.put(q.sp++, StackEntry{c[3], d[3]});
            if (d[2] < INFINITY) stk.put(q.sp++, StackEntry{c[2], d[2]});
            if (d[1] < INFINITY) stk.put(q.sp++, StackEntry{c[1], d[1]});
            query_enter(q, c[0]);
        }
    } else {
        float dl, dr;
        int32_t cl, cr;
        const int32_t ref = q.cur;
        if constexpr (DIM == 3) {
            const NodeF3 n = ref < n_cached ? static_cast<const NodeF3 *>(cache)[ref] : m.nf3[ref];
            dl = box_d2f(n.l_lox, n.l_loy, n.l_loz, n.l_hix, n.l_hiy, n.l_hiz, px, py, pz);
            dr = box_d2f(n.r_lox, n.r_loy, n.r_loz, n.r_hix, n.r_hiy, n.r_hiz, px, py, pz);
            cl = n.left; cr = n.right;
        } else {
            const NodeF2 n = ref < n_cached ? static_cast<const NodeF2 *>(cache)[ref] : m.nf2[ref];
            dl = box_d2f2(n.l_lox, n.l_loy, n.l_hix, n.l_hiy, px, py);
            dr = box_d2f2(n.r_lox, n.r_loy, n.r_hix, n.r_hiy, px, py);
            cl = n.left; cr = n.right;
        }
        did |= kQsNode;
        const float thr = q.best * q.cf;
        const bool hl = dl < thr, hr = dr < thr;
        if (!hl && dl < q.best) q.lb = fminf(q.lb, dl);
        if (!hr && dr < q.best) q.lb = fminf(q.lb, dr);
        pop = false;
        if (hl && hr) {
            const bool left_first = dl <= dr;
            StackEntry far;
            far.ref = left_first ? cr : cl;
            far.d2 = left_first ? dr : dl;
            stk.put(q.sp++, far);
            query_enter(q, left_first ? cl : cr);
        } else if (hl) {
            query_enter(q, cl);
        } else if (hr) {
            query_enter(q, cr);
        } else {
            pop = true;
        }
    }
    if (pop) {
        while (q.sp > 0) {
            const StackEntry e = stk.get(--q.sp);
            if (e.d2 < q.best * q.cf) {
                query_enter(q, e.ref);
                return did | kQsMore;
            }
            if (e.d2 < q.best) q.lb = fminf(q.lb, e.d2);
        }
        return did;
    }
    return did | kQsMore;
}

}  // namespace wosx
