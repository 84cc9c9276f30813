// ref_driver.cpp -- C entry points over the reference's own compiled core.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
// the reference sources src/{bvh,wos,mesh,inside,npq,prng}.cpp where they lie
// in the reference repository, into oracle/_ref/libwosref.so.  It is used to
// pin the C oracle bit-for-bit, to write the golden fixtures under
// tests/golden/, and as the "reference" CPU baseline of bench.py.
//
// The equation functors of src/equations/*.cpp sit in translation units that
// include the MPI-IO HDF5 writer (absent from this image's serial HDF5), so
// the functor bodies are restated here from laplace.cpp:12-43,
// poisson.cpp:15-118 and screened_poisson.cpp:40-141 and driven by the
// descriptor of include/wos_hip.h; everything they call (wos_run,
// wos_sample, bvh_npq, eberly, classify_point, poisson_green, screened,
// bessel, PRNG) is the reference's own code, and so is the Gaussian source
// proposal of GreenSourceMIS (apps/welding/include/welding/
// gaussian_heat_source.hpp).
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "wos/bvh.hpp"
#include "wos/fastmath.hpp"
#include "wos/grid.hpp"
#include "wos/hash.hpp"
#include "wos/inside.hpp"
#include "wos/mesh.hpp"
#include "wos/npq.hpp"
#include "wos/poisson.hpp"
#include "wos/prng.hpp"
#include "wos/screened.hpp"
#include "wos/wos.hpp"
#include "welding/field_sampler.hpp"
#include "welding/gaussian_heat_source.hpp"

#include "../include/wos_hip.h"

using namespace wos;

namespace {

double src_value(const wos_eq_desc &e, const double *x, int dim) {
    double acc = 0.0;
    for (int k = 0; k < dim; ++k) {
        const double d = x[k] - e.src_center[k];
        acc = k == 0 ? d * d : acc + d * d;
    }
    switch (e.source_kind) {
    case WOS_SRC_DISK_SQ: return acc <= e.src_radius * e.src_radius ? e.src_value : 0.0;
    case WOS_SRC_DISK_SQRT: return std::sqrt(acc) <= e.src_radius ? e.src_value : 0.0;
    case WOS_SRC_CONSTANT: return e.src_value;
    case WOS_SRC_GAUSSIAN: return e.src_value * std::exp(-acc / (e.src_radius * e.src_radius));
    default: return 0.0;
    }
}

template<bool Src, bool Scr>
struct Desc2D {
    static constexpr bool has_source = Src;
    static constexpr bool has_screening = Scr;
    static constexpr bool has_green_source = true;
    const wos_eq_desc *e;
    SourceMode source_mode;
    double alpha;
    // GreenSourceMIS: the Gaussian's own density on the proposal box, the
    // welding app's finite-plate source (level1_stationary_heat.cpp:180-190)
    const welding::GaussianHeatSource2D *proposal;
    // WOS_SRC_FIELD: the reference's FieldSampler2D over the previous field
    const welding::FieldSampler2D *field;

    Point2D sample_source_proposal(PRNG &rng) const {
        return proposal->sample_spatial_distribution(rng);
    }
    double source_proposal_pdf(Point2D p) const { return proposal->spatial_probability_density(p); }
    double source_mis_green_probability() const { return e->mis_green_probability; }

    bool source_may_intersect(Sphere2D s) const {
        if (e->source_kind != WOS_SRC_DISK_SQ && e->source_kind != WOS_SRC_DISK_SQRT) return true;
        const double dx = s.centre.x - e->src_center[0];
        const double dy = s.centre.y - e->src_center[1];
        const double reach = s.radius + e->src_radius;
        return dx * dx + dy * dy <= reach * reach;
    }
    double source(Point2D p) const {
        if (e->source_kind == WOS_SRC_FIELD) return field_value(p);
        const double x[2] = {p.x, p.y};
        return src_value(*e, x, 2);
    }
    // the welding sources (level0_cooling.cpp:194-196,
    // level2_transient_heat.cpp:255-264, level3_moving_heat.cpp:189-194)
    double field_value(Point2D p) const {
        const wos_field_source &f = *e->field_source;
        double history = 0.0;
        if (field) {
            const double v = field->sample(p);
            history = f.history_scale * (f.clamp_nonnegative ? std::max(0.0, v) : v);
        }
        double heat = 0.0;
        if (f.heat_scale != 0.0 && f.n_heat_terms > 0) {
            double sum = 0.0;
            for (int c = 0; c < f.n_heat_terms; ++c) {
                const wos_heat_term &h = f.heat[c];
                const double dx = (p.x - h.cx) / h.sx, dy = (p.y - h.cy) / h.sy;
                sum += h.a * std::exp(-0.5 * (dx * dx + dy * dy)) / h.b;
            }
            heat = f.heat_scale * sum / f.heat_divisor;
        }
        return history + heat;
    }
    double boundary(Point2D p, int bid) const {
        switch (e->boundary_kind) {
        case WOS_BND_X2Y2_ID0: return bid == 0 ? p.x * p.x - p.y * p.y : 0.0;
        case WOS_BND_ID_CONST: return (bid >= 0 && bid < e->n_bnd_values) ? e->bnd_values[bid] : 0.0;
        case WOS_BND_SIN3X_COS2Y: return std::sin(3.0 * p.x) * std::cos(2.0 * p.y);
        case WOS_BND_ONE: return 1.0;
        default: return 0.0;
        }
    }
    double green(Sphere2D s, Point2D x, Point2D y) const {
        const double r = dist(x, y);
        return Scr ? screened::green_2d(alpha, s.radius, r) : poisson_green::green_2d(s.radius, r);
    }
    // helmholtz.cpp:36-41 with k = alpha, else screened::weight_2d
    double screening_factor(double radius) const {
        if (e->kind == WOS_EQ_HELMHOLTZ) {
            const double prod = radius * alpha;
            if (prod < 1e-8) return 1.0;
            return 1.0 / bessel_J0(prod);
        }
        return screened::weight_2d(alpha, radius);
    }
    double green_mass(double radius) const {
        return Scr ? screened::mass_2d(alpha, radius) : poisson_green::mass_2d(radius);
    }
    Point2D sample_green(Sphere2D s, PRNG &rng) const {
        const double radius = Scr ? screened::sample_radius_2d(alpha, s.radius, rng)
                                  : poisson_green::sample_radius_2d(s.radius, rng);
        const double angle = 2.0 * poisson_green::pi * rng.unit();
        return Point2D{s.centre.x + radius * std::cos(angle), s.centre.y + radius * std::sin(angle)};
    }
};

template<bool Src, bool Scr>
struct Desc3D {
    static constexpr bool has_source = Src;
    static constexpr bool has_screening = Scr;
    static constexpr bool has_green_source = true;
    const wos_eq_desc *e;
    SourceMode source_mode;
    double alpha;

    double source(Point3D p) const { const double x[3] = {p.x, p.y, p.z}; return src_value(*e, x, 3); }
    double boundary(Point3D p) const {
        switch (e->boundary_kind) {
        case WOS_BND_Y5: {
            double r_sq = p.x * p.x + p.y * p.y;
            double y50 = p.z * (8 * p.z * p.z * p.z * p.z - 40 * p.z * p.z * r_sq + 15 * r_sq * r_sq);
            double y5m3 = (8 * p.z * p.z - r_sq) * p.y * (3 * p.x * p.x - p.y * p.y);
            return y50 + 4.0 * y5m3;
        }
        case WOS_BND_SIN3X_COS2Y: return std::sin(3.0 * p.x) * std::cos(2.0 * p.y);
        case WOS_BND_ONE: return 1.0;
        default: return 0.0;
        }
    }
    double green(Sphere3D s, Point3D x, Point3D y) const {
        const double r = dist(x, y);
        return Scr ? screened::green_3d(alpha, s.radius, r) : poisson_green::green_3d(s.radius, r);
    }
    // helmholtz.cpp:70-75 with k = alpha, else screened::weight_3d
    double screening_factor(double radius) const {
        if (e->kind == WOS_EQ_HELMHOLTZ) {
            const double prod = radius * alpha;
            if (prod < 1e-8) return 1.0;
            return prod / std::sin(prod);
        }
        return screened::weight_3d(alpha, radius);
    }
    double green_mass(double radius) const {
        return Scr ? screened::mass_3d(alpha, radius) : poisson_green::mass_3d(radius);
    }
    Point3D sample_green(Sphere3D s, PRNG &rng) const {
        const double radius = Scr ? screened::sample_radius_3d(alpha, s.radius, rng)
                                  : poisson_green::sample_radius_3d(s.radius, rng);
        const double z = 2.0 * rng.unit() - 1.0;
        const double angle = 2.0 * poisson_green::pi * rng.unit();
        const double xy = std::sqrt(1.0 - z * z);
        return Point3D{s.centre.x + radius * xy * std::cos(angle),
                       s.centre.y + radius * xy * std::sin(angle), s.centre.z + radius * z};
    }
};

struct Handle {
    int dim = 0;
    Mesh<2> m2;
    Mesh<3> m3;
    std::unique_ptr<BVH<2>> b2;
    std::unique_ptr<BVH<3>> b3;
};

template<int N, typename Eq>
WoSResult run_point(const BVH<N> &bvh, Point<N> p0, const Eq &eq, int n_walks, double eps,
                    int max_steps, std::uint64_t ps, double *vals, int *steps) {
    auto start = get_start_point(bvh, p0);
    PRNG walk_rng(splitmix64(ps ^ 0x13198A2E03707344ULL));
    if (!vals) {
        if constexpr (N == 2) {
            PRNG source_rng(splitmix64(ps ^ 0xA4093822299F31D0ULL));
            return wos_run(bvh, p0, start, eq, n_walks, eps, walk_rng, source_rng, max_steps);
        } else {
            return wos_run(bvh, p0, start, eq, n_walks, eps, walk_rng, walk_rng, max_steps);
        }
    }
    // per-walk values: the same loop as wos_run, exposing each sample
    PRNG source_rng(splitmix64(ps ^ 0xA4093822299F31D0ULL));
    PRNG &src = (N == 2) ? source_rng : walk_rng;
    for (int i = 0; i < n_walks; ++i) {
        SampleResult s = wos_sample(bvh, p0, start, eq, eps, walk_rng, src, max_steps);
        vals[i] = s.value;
        steps[i] = s.steps;
    }
    return WoSResult{0, 0, 0, 0, 0};
}

template<int N, typename Eq>
void dispatch_points(Handle *h, const Eq &eq, const double *pts, const int64_t *seed_index,
                     int64_t n, int n_walks, double eps, int max_steps, int max_ray_attempts,
                     std::uint64_t seed, int n_threads, double *mean, double *var, double *se,
                     double *msteps, uint8_t *loc, int32_t *hits, int32_t *amb) {
    const BVH<N> *bvhp;
    if constexpr (N == 2) bvhp = h->b2.get();
    else bvhp = h->b3.get();
    const BVH<N> &bvh = *bvhp;
    double g[6];
    if constexpr (N == 2) mesh_bbox(h->m2, &g[0], &g[1], &g[2], &g[3], &g[4], &g[5]);
    else mesh_bbox(h->m3, &g[0], &g[1], &g[2], &g[3], &g[4], &g[5]);
    const double tol = 1e-10 * std::max({1.0, g[1] - g[0], g[3] - g[2], g[5] - g[4]});
    std::atomic<int64_t> next{0};
    auto body = [&]() {
        for (;;) {
            const int64_t i = next.fetch_add(1);
            if (i >= n) break;
            const std::uint64_t lin = seed_index ? (std::uint64_t)seed_index[i] : (std::uint64_t)i;
            const std::uint64_t ps = splitmix64(seed ^ splitmix64(lin));
            Point<N> p0;
            if constexpr (N == 2) p0 = Point2D{pts[2 * i], pts[2 * i + 1]};
            else p0 = Point3D{pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
            auto start = get_start_point(bvh, p0);
            int code, retries = 0;
            if constexpr (N == 2) {
                code = (int)classify_point(h->m2, p0, start.radius, tol);
            } else {
                PRNG inside_rng(splitmix64(ps ^ 0x243F6A8885A308D3ULL));
                auto c = classify_point(bvh, p0, start.radius, tol, max_ray_attempts, inside_rng);
                code = (int)c.location;
                retries = c.ambiguous_ray_retries;
            }
            WoSResult r{NAN, NAN, NAN, NAN, 0};
            if (code == 2) {
                if constexpr (N == 2) r = WoSResult{eq.boundary(start.nearest, start.boundary_id), 0, 0, 0, 0};
                else r = WoSResult{eq.boundary(start.nearest), 0, 0, 0, 0};
            } else if (code == 1) {
                r = run_point<N>(bvh, p0, eq, n_walks, eps, max_steps, ps, nullptr, nullptr);
            }
            mean[i] = r.mean;
            if (var) var[i] = r.variance;
            if (se) se[i] = r.standard_error;
            if (msteps) msteps[i] = r.mean_steps;
            if (loc) loc[i] = (uint8_t)code;
            if (hits) hits[i] = r.max_steps_hits;
            if (amb) amb[i] = retries;
        }
    };
    if (n_threads <= 1) { body(); return; }
    std::vector<std::thread> th;
    for (int t = 0; t < n_threads; ++t) th.emplace_back(body);
    for (auto &t : th) t.join();
}

template<int N, typename F>
void with_eq(const wos_eq_desc *e, F &&f) {
    const SourceMode mode = static_cast<SourceMode>(e->source_mode);
    const bool src = e->source_kind != WOS_SRC_NONE;
    const bool scr = e->kind == WOS_EQ_SCREENED || e->kind == WOS_EQ_HELMHOLTZ;
    if constexpr (N == 2) {
        // the proposal of a Gaussian source: sigma = R / sqrt 2 per axis,
        // unit power, the descriptor's box (+-inf limits without one)
        std::unique_ptr<welding::GaussianHeatSource2D> hs;
        if (src && mode == SourceMode::GreenSourceMIS && e->source_kind == WOS_SRC_GAUSSIAN) {
            const double *b = e->proposal_box;
            const bool boxed = !(b[0] == 0.0 && b[1] == 0.0 && b[2] == 0.0 && b[3] == 0.0);
            const double inf = INFINITY;
            const double sigma = e->src_radius / std::sqrt(2.0);
            hs = std::make_unique<welding::GaussianHeatSource2D>(
                boxed ? b[0] : -inf, boxed ? b[1] : inf, boxed ? b[2] : -inf, boxed ? b[3] : inf,
                1.0, 1.0, 1.0, e->src_center[0], e->src_center[1], sigma, sigma);
        }
        const welding::GaussianHeatSource2D *pr = hs.get();
        // the previous field of a WOS_SRC_FIELD source
        std::unique_ptr<Grid> fg;
        std::vector<double> fv;
        std::unique_ptr<welding::FieldSampler2D> fs;
        const wos_field_source *fd = e->source_kind == WOS_SRC_FIELD ? e->field_source : nullptr;
        if (fd && fd->values) {
            fg = std::make_unique<Grid>(Grid{fd->nx, fd->ny, 1, fd->bounds[0], fd->bounds[1],
                                             fd->bounds[2], fd->bounds[3], 0.0, 0.0});
            fv.assign(fd->values, fd->values + (size_t)fd->nx * fd->ny);
            fs = std::make_unique<welding::FieldSampler2D>(
                *fg, fv, fd->interpolation == WOS_INTERP_CUBIC ? welding::FieldInterpolation2D::Cubic
                                                               : welding::FieldInterpolation2D::Bilinear);
        }
        const welding::FieldSampler2D *fp = fs.get();
        if (src && scr) f(Desc2D<true, true>{e, mode, e->alpha, pr, fp});
        else if (src) f(Desc2D<true, false>{e, mode, e->alpha, pr, fp});
        else if (scr) f(Desc2D<false, true>{e, mode, e->alpha, pr, fp});
        else f(Desc2D<false, false>{e, mode, e->alpha, pr, fp});
    } else {
        if (src && scr) f(Desc3D<true, true>{e, mode, e->alpha});
        else if (src) f(Desc3D<true, false>{e, mode, e->alpha});
        else if (scr) f(Desc3D<false, true>{e, mode, e->alpha});
        else f(Desc3D<false, false>{e, mode, e->alpha});
    }
}

}  // namespace

extern "C" {

uint64_t ref_splitmix64(uint64_t x) { return splitmix64(x); }

void ref_prng_draws(uint64_t seed, int n, uint64_t *out) {
    PRNG rng(seed);
    for (int i = 0; i < n; ++i) out[i] = rng.u64();
}

double ref_npq_seg(const double *s0, const double *s1, const double *p, double *nearest) {
    Point2D np;
    const double d = npq_seg(Point2D{s0[0], s0[1]}, Point2D{s1[0], s1[1]}, Point2D{p[0], p[1]}, &np);
    nearest[0] = np.x; nearest[1] = np.y;
    return d;
}

double ref_eberly(const double *p, const double *v0, const double *v1, const double *v2,
                  double *nearest) {
    // the cache fill of src/bvh.cpp:190-205
    Vec3D e1{v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
    Vec3D e2{v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
    double a = dot(e1, e1), b = dot(e1, e2), c = dot(e2, e2);
    EberlyCache cache{Point3D{v0[0], v0[1], v0[2]}, e1, e2, a, b, c, a * c - b * b, a - 2 * b + c};
    Point3D np;
    const double d = eberly(Point3D{p[0], p[1], p[2]}, &np, &cache);
    nearest[0] = np.x; nearest[1] = np.y; nearest[2] = np.z;
    return d;
}

int ref_moller_trumbore(const double *p, const double *ray, const double *a, const double *b,
                        const double *c) {
    return (int)moller_trumbore(Point3D{p[0], p[1], p[2]}, Vec3D{ray[0], ray[1], ray[2]},
                                Point3D{a[0], a[1], a[2]}, Point3D{b[0], b[1], b[2]},
                                Point3D{c[0], c[1], c[2]});
}

void *ref_mesh_create(int dim, const double *verts, int n_verts, const int *prims, int n_prims,
                      const int *bids) {
    auto *h = new Handle;
    h->dim = dim;
    if (dim == 2) {
        h->m2.verts.resize(n_verts);
        for (int i = 0; i < n_verts; ++i) h->m2.verts[i] = Point2D{verts[2 * i], verts[2 * i + 1]};
        h->m2.prims.assign(prims, prims + 2 * n_prims);
        h->m2.boundary_ids.assign(n_prims, 0);
        if (bids) h->m2.boundary_ids.assign(bids, bids + n_prims);
        h->b2 = build_bvh<2>(h->m2);
    } else {
        h->m3.verts.resize(n_verts);
        for (int i = 0; i < n_verts; ++i)
            h->m3.verts[i] = Point3D{verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]};
        h->m3.prims.assign(prims, prims + 3 * n_prims);
        h->b3 = build_bvh<3>(h->m3);
    }
    return h;
}

void ref_mesh_free(void *p) { delete static_cast<Handle *>(p); }

int ref_bvh_n_nodes(void *p) {
    auto *h = static_cast<Handle *>(p);
    return h->dim == 2 ? (int)h->b2->nodes.size() : (int)h->b3->nodes.size();
}

void ref_bvh_export(void *p, int *prim_count, int *idx, double *bmin, double *bmax, int *leaf_prims) {
    auto *h = static_cast<Handle *>(p);
    if (h->dim == 2) {
        for (size_t i = 0; i < h->b2->nodes.size(); ++i) {
            const auto &n = h->b2->nodes[i];
            prim_count[i] = n.prim_count; idx[i] = n.idx;
            bmin[3 * i] = n.bbox.pmin.x; bmin[3 * i + 1] = n.bbox.pmin.y; bmin[3 * i + 2] = 0;
            bmax[3 * i] = n.bbox.pmax.x; bmax[3 * i + 1] = n.bbox.pmax.y; bmax[3 * i + 2] = 0;
        }
        std::memcpy(leaf_prims, h->b2->prims.data(), sizeof(int) * h->b2->prims.size());
    } else {
        for (size_t i = 0; i < h->b3->nodes.size(); ++i) {
            const auto &n = h->b3->nodes[i];
            prim_count[i] = n.prim_count; idx[i] = n.idx;
            bmin[3 * i] = n.bbox.pmin.x; bmin[3 * i + 1] = n.bbox.pmin.y; bmin[3 * i + 2] = n.bbox.pmin.z;
            bmax[3 * i] = n.bbox.pmax.x; bmax[3 * i + 1] = n.bbox.pmax.y; bmax[3 * i + 2] = n.bbox.pmax.z;
        }
        std::memcpy(leaf_prims, h->b3->prims.data(), sizeof(int) * h->b3->prims.size());
    }
}

// nearest point query: returns distance, writes nearest point and (2D) prim/boundary id
double ref_bvh_npq(void *p, const double *q, double *nearest, int *prim, int *bid) {
    auto *h = static_cast<Handle *>(p);
    if (h->dim == 2) {
        auto r = bvh_npq(*h->b2, Point2D{q[0], q[1]});
        nearest[0] = r.point.x; nearest[1] = r.point.y;
        *prim = r.primitive_id; *bid = r.boundary_id;
        return r.distance;
    }
    Point3D np;
    const double d = bvh_npq(*h->b3, Point3D{q[0], q[1], q[2]}, &np);
    nearest[0] = np.x; nearest[1] = np.y; nearest[2] = np.z;
    *prim = -1; *bid = -1;
    return d;
}

int ref_winding_number(void *p, const double *q) {
    return winding_number(static_cast<Handle *>(p)->m2, Point2D{q[0], q[1]});
}

// per-walk sample values of one point (2D: separate walk/source streams; 3D shared)
void ref_walks(void *p, const wos_eq_desc *e, const double *p0, int n_walks, double eps,
               int max_steps, uint64_t point_seed, double *vals, int *steps) {
    auto *h = static_cast<Handle *>(p);
    if (h->dim == 2) {
        with_eq<2>(e, [&](const auto &eq) {
            run_point<2>(*h->b2, Point2D{p0[0], p0[1]}, eq, n_walks, eps, max_steps, point_seed, vals, steps);
        });
    } else {
        with_eq<3>(e, [&](const auto &eq) {
            run_point<3>(*h->b3, Point3D{p0[0], p0[1], p0[2]}, eq, n_walks, eps, max_steps, point_seed, vals, steps);
        });
    }
}

// the sweep body of runner.hpp:157-235 over a point list, n_threads workers
void ref_solve_points(void *p, const wos_eq_desc *e, const double *pts, const int64_t *seed_index,
                      int64_t n, int n_walks, double eps, int max_steps, int max_ray_attempts,
                      uint64_t seed, int n_threads, double *mean, double *var, double *se,
                      double *msteps, uint8_t *loc, int32_t *hits, int32_t *amb) {
    auto *h = static_cast<Handle *>(p);
    if (h->dim == 2) {
        with_eq<2>(e, [&](const auto &eq) {
            dispatch_points<2>(h, eq, pts, seed_index, n, n_walks, eps, max_steps, max_ray_attempts,
                               seed, n_threads, mean, var, se, msteps, loc, hits, amb);
        });
    } else {
        with_eq<3>(e, [&](const auto &eq) {
            dispatch_points<3>(h, eq, pts, seed_index, n, n_walks, eps, max_steps, max_ray_attempts,
                               seed, n_threads, mean, var, se, msteps, loc, hits, amb);
        });
    }
}

}  // extern "C"
