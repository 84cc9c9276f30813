// wos_engine.hip -- HIP kernels (gfx950) and the C ABI of include/wos_hip.h.
//
// Pipeline of one solve (replacing the body of run<N,Eq>,
// include/wos/runner.hpp:82-238):
//   1. k_classify: per point, FP64 start query (reference traversal), the
//      reference boundary / winding / ray-vote classification, the walk-stream
//      key and g at the nearest point;
//   2. compaction of the inside points (on the device for a one-round grid
//      solve);
//   3. the walker: k_walk_fp32 (FP32 queries on the distance grid), k_walk_fp64
//      (the reference's FP64 search, resumable) or k_walk_ref (the reference's
//      per-point streams), all on persistent lanes refilled from a pool,
//      writing one FP64 value per walk into a point-major staging buffer;
//   4. k_welford: per point, the reference's sequential Welford update in walk
//      order (include/wos/wos.hpp:169-180), continued across rounds (FP64
//      walkers); k_merge_wave: the same update on 64 lanes per point and a
//      fixed-order pairwise combination (FP32 walker).
// Results are bitwise reproducible: a walk's draws depend only on (point seed,
// walk, step), and the merge order is fixed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/wos_hip.h"
#include "bvh_build.hpp"
#include "wos_common.hpp"
#include "wos_device.hpp"

using namespace wosx;

namespace {

constexpr int kBlock = 256;
constexpr int kMaxParts = 64;  // largest xcd_parts

// --------------------------------------------------------------------------
// walk-step helpers (FP64): src/wos.cpp:9-56, include/wos/wos.hpp:54-138
// --------------------------------------------------------------------------
// The samplers take their uniforms from an RNG object with unit() and
// unit_open(): the counter streams of one walk step (StepStreamD/F,
// wos_common.hpp) or a reference xoshiro stream (XoRef).  RNG::radius_first
// orders the draws of a uniform ball point: the counter streams draw its
// radius first (so an FP32 constant-source step needs one Philox block: the
// direction's and the radius's slots share it), the reference draws the
// direction first (src/wos.cpp:33-56).

// include/wos/poisson.hpp:27-30, :46-50
template <typename T, typename RNG>
__device__ __forceinline__ T poisson_radius_2d(T R, RNG &rng) {
    const T u = rng.unit_open();
    const T v = rng.unit_open();
    return R * Fx<T>::sqrt(u * v);
}
template <typename T, typename RNG>
__device__ __forceinline__ T poisson_radius_3d(T R, RNG &rng) {
    const T u1 = rng.unit_open(), u2 = rng.unit_open();
    const T a = -Fx<T>::log(u1 * u2);
    const T v1 = rng.unit_open(), v2 = rng.unit_open();
    const T b = -Fx<T>::log(v1 * v2);
    return Fx<T>::div(R * a, a + b);
}
// include/wos/screened.hpp:97-101
template <typename T, typename RNG>
__device__ __forceinline__ T normal_draw(RNG &rng) {
    const T radius = sqrt((T)-2 * log(rng.unit_open()));
    const T angle = (T)2 * (T)kPi * rng.unit();
    return radius * cos(angle);
}
template <typename T> __device__ __forceinline__ T clamp01(T v) {
    return v < (T)0 ? (T)0 : (v > (T)1 ? (T)1 : v);
}
// The reference's rejection loops (include/wos/screened.hpp:105-156) run
// until a proposal is accepted.  Here a loop stops after kRadiusMaxTrials
// and returns NaN: the walk's value becomes NaN, the point's mean shows it
// and k_welford counts the walk into the summary's failed_walks, instead of
// a biased radius being used silently.  (The acceptance rate is bounded away
// from zero for every alpha R, so no walk is expected to stop here.)
constexpr int kRadiusMaxTrials = 100000;

// include/wos/screened.hpp:105-129
// (the sphere's log K0(X) and log I0(X) are taken once, outside the
// rejection loop: the same values in the same expression)
template <typename T, typename RNG>
__device__ __forceinline__ T screened_radius_2d(T alpha, T R, RNG &rng) {
    const T X = alpha * R;
    const T lk_X = X < (T)2 ? (T)0 : log_k0<T>(X);
    const T li_X = X < (T)2 ? (T)0 : log_i0<T>(X);
    for (int guard = 0; guard < kRadiusMaxTrials; ++guard) {
        T r, acceptance;
        if (X < (T)2) {
            r = poisson_radius_2d<T>(R, rng);
            const T proposal = poisson_green_2d<T>(R, r);
            acceptance = screened_green_2d<T>(alpha, R, r) / proposal;
        } else {
            const T time = -log(rng.unit_open()) / (alpha * alpha);
            r = sqrt((T)-4 * time * log(rng.unit_open()));
            if (r >= R) continue;
            const T x = alpha * r;
            const T log_ratio = log_i0<T>(x) + lk_X - li_X - log_k0<T>(x);
            acceptance = -expm1(log_ratio < (T)0 ? log_ratio : (T)0);
        }
        if (rng.unit() < clamp01<T>(acceptance)) return r;
    }
    return (T)NAN;
}
// include/wos/screened.hpp:131-156
template <typename T, typename RNG>
__device__ __forceinline__ T screened_radius_3d(T alpha, T R, RNG &rng) {
    const T X = alpha * R;
    for (int guard = 0; guard < kRadiusMaxTrials; ++guard) {
        T r, acceptance;
        if (X < (T)2) {
            r = poisson_radius_3d<T>(R, rng);
            const T proposal = poisson_green_3d<T>(R, r);
            acceptance = screened_green_3d<T>(alpha, R, r) / proposal;
        } else {
            const T time = -log(rng.unit_open()) / (alpha * alpha);
            const T scale = sqrt((T)2 * time);
            const T x0 = scale * normal_draw<T>(rng);
            const T x1 = scale * normal_draw<T>(rng);
            const T x2 = scale * normal_draw<T>(rng);
            r = sqrt(x0 * x0 + x1 * x1 + x2 * x2);
            if (r >= R) continue;
            const T x = alpha * r;
            acceptance = (-expm1((T)-2 * (X - x))) / (-expm1((T)-2 * X));
        }
        if (rng.unit() < clamp01<T>(acceptance)) return r;
    }
    return (T)NAN;
}

// attempts of the truncated-Gaussian rejection sampler before it gives up
// (gaussian_heat_source.hpp:58); the term is then NaN, so the point's mean
// shows the failure
constexpr int kMisMaxAttempts = 1000000;

// Green-source mixture term of a 2D Gaussian source (include/wos/
// source_sampling.hpp:35-78, the weight applied as at wos.hpp:86-88).  A
// first draw picks the Green sample of the Green mode (probability beta) or
// a draw of the source's own density q: a normal of per-axis width sigma
// about the centre by Box-Muller, rejected until it falls in the truncation
// box (gaussian_heat_source.hpp:55-74).  A sample outside the open ball adds
// nothing; otherwise the term is weight G f / (beta G / mass + (1 - beta) q)
// with q = exp(-((x-cx)/sigma)^2/2 - ((y-cy)/sigma)^2/2) / norm inside the
// box, 0 outside (gaussian_heat_source.hpp:45-51, :109-113).
template <typename T, typename RNG, bool SCR = true>
__device__ double source_term_mis(const EqView &e, T px, T py, T R, T weight, RNG &rng) {
    const bool screened = SCR && e.kind == WOS_EQ_SCREENED;
    const T alpha = (T)e.alpha, beta = (T)e.mis_beta;
    const T cx = (T)e.cx, cy = (T)e.cy, sigma = (T)e.mis_sigma;
    T sx, sy;
    if (rng.unit() < beta) {
        const T radius = screened ? screened_radius_2d<T>(alpha, R, rng) : poisson_radius_2d<T>(R, rng);
        const T angle = (T)2 * (T)kPi * rng.unit();
        sx = px + radius * Fx<T>::cos(angle);
        sy = py + radius * Fx<T>::sin(angle);
    } else {
        for (int attempt = 0;; ++attempt) {
            if (attempt == kMisMaxAttempts) return NAN;
            const T magnitude = Fx<T>::sqrt((T)-2 * Fx<T>::log(rng.unit_open()));
            const T angle = (T)2 * (T)kPi * rng.unit();
            sx = cx + sigma * magnitude * Fx<T>::cos(angle);
            sy = cy + sigma * magnitude * Fx<T>::sin(angle);
            if (!e.mis_box || (sx >= (T)e.box[0] && sx <= (T)e.box[1] && sy >= (T)e.box[2] &&
                               sy <= (T)e.box[3]))
                break;
        }
    }
    const T dx = px - sx, dy = py - sy;
    const T dist = Fx<T>::sqrt(dx * dx + dy * dy);
    if (!(dist > (T)0 && dist < R)) return 0.0;
    const T mass = screened ? screened_mass_2d<T>(alpha, R) : R * R / (T)4;
    const T G = screened ? screened_green_2d<T>(alpha, R, dist) : poisson_green_2d<T>(R, dist);
    const T green_pdf = Fx<T>::div(G, mass);
    T q = (T)0;
    if (!e.mis_box || (sx >= (T)e.box[0] && sx <= (T)e.box[1] && sy >= (T)e.box[2] &&
                       sy <= (T)e.box[3])) {
        const T ux = Fx<T>::div(sx - cx, sigma), uy = Fx<T>::div(sy - cy, sigma);
        q = Fx<T>::div(Fx<T>::exp((T)-0.5 * (ux * ux + uy * uy)), (T)e.mis_norm);
    }
    const T mixture = beta * green_pdf + ((T)1 - beta) * q;
    if (!(mixture > (T)0) || !(mixture < (T)INFINITY)) return NAN;
    const T f = eq_source<T>(e, sx, sy, (T)0, 2);
    if (f == (T)0) return 0.0;
    return (double)(weight * Fx<T>::div(G * f, mixture));
}

// One source term of wos_sample (include/wos/wos.hpp:69-110): returns the
// value to add (already multiplied by the weight).  T is the geometry type;
// SX selects the source extensions compiled in: 0 none, 1 the GREEN_MIS
// branch, 2 the WOS_SRC_FIELD source.  Each goes only into the kernels
// launched for it, so the other kernels keep their registers.  SCR = false
// (the kernels for Laplace and Poisson) leaves the screened samplers out.
template <int DIM, typename T, typename RNG, int SX = 0, bool SCR = true>
__device__ __forceinline__ double source_term(const EqView &e, T px, T py, T pz, T R, T weight,
                                              RNG &rng) {
    const bool screened = SCR && e.kind == WOS_EQ_SCREENED;
    const T alpha = (T)e.alpha;
    if (DIM == 2 && !eq_source_may_intersect<T>(e, px, py, R)) return 0.0;
    if constexpr (SX == 1 && DIM == 2)
        if (e.source_mode == WOS_SOURCE_GREEN_MIS)
            return source_term_mis<T, RNG, SCR>(e, px, py, R, weight, rng);
    constexpr bool FLD = SX == 2;
    T sx, sy, sz = (T)0;
    if (e.source_mode == WOS_SOURCE_GREEN) {
        if (DIM == 2) {
            const T radius = screened ? screened_radius_2d<T>(alpha, R, rng) : poisson_radius_2d<T>(R, rng);
            const T angle = (T)2 * (T)kPi * rng.unit();
            sx = px + radius * Fx<T>::cos(angle);
            sy = py + radius * Fx<T>::sin(angle);
        } else {
            const T radius = screened ? screened_radius_3d<T>(alpha, R, rng) : poisson_radius_3d<T>(R, rng);
            const T z = (T)2 * rng.unit() - (T)1;
            const T angle = (T)2 * (T)kPi * rng.unit();
            const T xy = Fx<T>::sqrt((T)1 - z * z);
            sx = px + radius * xy * Fx<T>::cos(angle);
            sy = py + radius * xy * Fx<T>::sin(angle);
            sz = pz + radius * z;
        }
        const T f = eq_source<T, FLD>(e, sx, sy, sz, DIM);
        if (f == (T)0) return 0.0;
        T mass;
        if (DIM == 2) mass = screened ? screened_mass_2d<T>(alpha, R) : R * R / (T)4;
        else mass = screened ? screened_mass_3d<T>(alpha, R) : R * R / (T)6;
        return (double)(weight * mass * f);
    }
    // uniform ball point (src/wos.cpp:33-56)
    T dist;
    if (DIM == 2) {
        T theta, r;
        if constexpr (RNG::radius_first) {
            r = R * sqrt(rng.unit_open());
            theta = rng.unit() * (T)2 * (T)kPi;
        } else {
            theta = rng.unit() * (T)2 * (T)kPi;
            r = R * sqrt(rng.unit_open());
        }
        sx = px + r * cos(theta);
        sy = py + r * sin(theta);
        const T dx = px - sx, dy = py - sy;
        dist = sqrt(dx * dx + dy * dy);
    } else {
        T u, v, w;
        if constexpr (RNG::radius_first) {
            w = rng.unit_open();
            u = rng.unit();
            v = rng.unit();
        } else {
            u = rng.unit();
            v = rng.unit();
            w = rng.unit_open();
        }
        const T z = (T)2 * u - (T)1, phi = (T)2 * (T)kPi * v;
        const T surf = sqrt((T)1 - z * z), vol = R * cbrt(w);
        sx = px + vol * surf * cos(phi);
        sy = py + vol * surf * sin(phi);
        sz = pz + vol * z;
        const T dx = px - sx, dy = py - sy, dz = pz - sz;
        dist = sqrt(dx * dx + dy * dy + dz * dz);
    }
    const T f = eq_source<T, FLD>(e, sx, sy, sz, DIM);
    if (f == (T)0) return 0.0;
    T volume;
    if (DIM == 2) volume = (T)kPi * R * R;
    else { const T rc = R * R * R; volume = (T)4 * (T)kPi * rc / (T)3; }
    T G;
    if (DIM == 2) G = screened ? screened_green_2d<T>(alpha, R, dist) : poisson_green_2d<T>(R, dist);
    else G = screened ? screened_green_3d<T>(alpha, R, dist) : poisson_green_3d<T>(R, dist);
    return (double)(weight * volume * f * G);
}

// --------------------------------------------------------------------------
// kernel 1: start query + classification (FP64, bit-faithful)
// --------------------------------------------------------------------------
struct GridSpec {
    int nx, ny, nz;
    double b[6];
};

// the domain a classify launch works on
struct DomainView {
    int root64;        // FP64 root node
    int32_t root32;    // FP32 root reference
    int s0, s1;        // leaf-slot range (winding number)
    double tol;        // boundary tolerance 1e-10 max(1, extent) (runner.hpp:86-88)
    int64_t out_off;   // first output index of this domain's points
    float shrink_abs;  // absolute margin of FP32 radii in this domain
    int index;         // domain index
};

__device__ __forceinline__ double grid_coordinate(double lo, double hi, int index, int count) {
    return count > 1 ? lo + (hi - lo) * (double)index / (double)(count - 1) : lo;
}

template <int DIM>
__device__ __forceinline__ void point_coords(const GridSpec &g, const double *plist, int64_t i,
                                             double &x, double &y, double &z) {
    if (plist) {
        x = plist[i * DIM];
        y = plist[i * DIM + 1];
        if (DIM == 3) z = plist[i * DIM + 2];
    } else {
        const int gi = (int)(i / ((int64_t)g.ny * g.nz));
        const int gj = (int)((i / g.nz) % g.ny);
        const int gk = (int)(i % g.nz);
        x = grid_coordinate(g.b[0], g.b[1], gi, g.nx);
        y = grid_coordinate(g.b[2], g.b[3], gj, g.ny);
        z = grid_coordinate(g.b[4], g.b[5], gk, g.nz);
    }
}

// Location code, walk keys and start record of point i from its start query
// q (src/inside.cpp:199-260, runner.hpp:182-221).  first_rays (3D): -1, or
// the first three ray casts made beforehand, bit k the hit parity of ray k
// and bit 3 + k set when ray k was ambiguous.  first_rays (2D): -1, or the
// inside flag (winding number != 0) found beforehand.
template <int DIM>
__device__ __forceinline__ void classify_tail(const MeshView &m, const EqView &e,
                                              const DomainView &dom, int64_t i, double x,
                                              double y, double z, const NpqD &q, uint64_t ps,
                                              int first_rays, int max_ray_attempts, int *stk,
                                              int stride, uint8_t *loc, StartInfo *starts,
                                              int32_t *amb, double *bvalue) {
    const double r = sqrt(q.d2);
    int code;
    int retries = 0;
    if (r <= dom.tol) {
        code = 2;
    } else if (DIM == 2) {
        code = first_rays >= 0 ? first_rays : (winding_number_d(m, dom.s0, dom.s1, x, y) != 0 ? 1 : 0);
    } else {
        // src/inside.cpp:228-260 with the inside stream of runner.hpp:182-183;
        // when none of the first three casts was ambiguous they decide
        // (three valid votes, no retries), else the loop runs from the start
        Xoshiro rng(splitmix64(ps ^ kInsideStream));
        int valid = 0, votes = 0;
        if (first_rays >= 0 && (first_rays >> 3) == 0) {
            valid = 3;
            votes = (first_rays & 1) + ((first_rays >> 1) & 1) + ((first_rays >> 2) & 1);
        }
        for (int attempts = 0; attempts < max_ray_attempts && valid < 3; ++attempts) {
            const double u = rng.unit(), v = rng.unit();
            const double zz = 2 * u - 1;
            const double phi = 2 * kPi * v;
            const double rr = sqrt(1 - zz * zz);
            bool ambiguous;
            const int hits = ray_cast_d(m, dom.root64, x, y, z, rr * cos(phi), rr * sin(phi), zz,
                                        ambiguous, stk, stride);
            if (ambiguous) { ++retries; continue; }
            votes += hits & 1;
            ++valid;
        }
        code = valid < 3 ? 3 : (votes > 1 ? 1 : 0);
    }
    const double gval = eq_boundary(e, m, q.px, q.py, q.pz, q.prim, q.s, q.t);
    const uint64_t key = splitmix64(ps ^ kWalkStream);
    StartInfo S;
    S.x = x; S.y = y; S.z = z; S.r = r;
    S.nx = q.px; S.ny = q.py; S.nz = q.pz;
    S.g = gval; S.s = q.s; S.t = q.t;
    S.prim = q.prim;
    S.slot = q.slot;
    S.root64 = dom.root64;
    S.root32 = dom.root32;
    S.shrink_abs = dom.shrink_abs;
    S.domain = dom.index;
    S.pad = 0;
    S.bid = (DIM == 2 && q.prim >= 0) ? m.prim_bid[q.prim] : -1;
    S.key0 = (uint32_t)key;
    S.key1 = (uint32_t)(key >> 32);
    const uint64_t skey = splitmix64(ps ^ kSourceStream);
    S.skey0 = (uint32_t)skey;
    S.skey1 = (uint32_t)(skey >> 32);
    S.pad1 = S.pad2 = 0u;
    const int64_t o = dom.out_off + i;
    starts[o] = S;
    loc[o] = (uint8_t)code;
    amb[o] = retries;
    bvalue[o] = gval;
}

template <int DIM>
__global__ __launch_bounds__(kBlock) void k_classify(MeshView m, EqView e, GridSpec g,
                                                     DomainView dom, const double *plist,
                                                     const int64_t *sidx, int64_t n, uint64_t seed,
                                                     int max_ray_attempts, uint8_t *loc,
                                                     StartInfo *starts, int32_t *amb,
                                                     double *bvalue) {
    extern __shared__ int lds_i[];
    int *stk = lds_i + threadIdx.x;
    const int stride = blockDim.x;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x = 0.0, y = 0.0, z = 0.0;
    point_coords<DIM>(g, plist, i, x, y, z);
    const uint64_t ps = point_seed(seed, sidx ? (uint64_t)sidx[i] : (uint64_t)i);
    NpqD q;
    if (DIM == 2) q = npq_ref_2d(m, dom.root64, x, y, stk, stride);
    else q = npq_ref_3d(m, dom.root64, x, y, z, stk, stride);
    classify_tail<DIM>(m, e, dom, i, x, y, z, q, ps, -1, max_ray_attempts, stk, stride, loc, starts,
                       amb, bvalue);
}

// The start query on one of S = 2^D subtrees: subtree j is reached by
// taking, at each of the top D levels, the near child (bit 0) or the far
// child (bit 1) in the reference's near-first order, most significant bit
// first, so j is the subtree's rank in the reference's traversal.  A leaf
// above depth D is searched by the lane whose remaining bits are 0.  Within
// the subtree the search is the reference's, so it returns the subtree's
// first primitive (in that order) at its least distance; over the subtrees
// the least distance with the lowest j is the reference's own result.
template <int S>
__device__ __forceinline__ NpqD npq_ref_3d_sub(const MeshView &m, int root, double px, double py,
                                               double pz, int j, int *stk, int stride,
                                               int budget = 0) {
    constexpr int D = S == 64 ? 6 : S == 32 ? 5 : S == 16 ? 4 : S == 8 ? 3 : S == 4 ? 2 : S == 2 ? 1 : 0;
    static_assert((1 << D) == S, "S must be a power of two up to 64");
    int curr = root;
    for (int l = 0; l < D; ++l) {
        const NodeD3 n = m.nd3[curr];
        if (n.count > 0) {
            if ((j & ((1 << (D - l)) - 1)) != 0) {
                NpqD r;
                r.d2 = INFINITY; r.px = r.py = r.pz = 0.0; r.s = r.t = 0.0; r.prim = -1; r.slot = -1;
                return r;
            }
            break;
        }
        const int left = curr + 1, right = n.idx;
        const bool left_near = box_d2_3(m.nd3[left], px, py, pz) <= box_d2_3(m.nd3[right], px, py, pz);
        const bool far = ((j >> (D - 1 - l)) & 1) != 0;
        curr = far == left_near ? right : left;
    }
    return npq_ref_3d(m, curr, px, py, pz, stk, stride, budget);
}

// npq_ref_3d_sub for the 2D tree: the same near/far descent over the top D
// levels (near = the child with the smaller box distance, ties to the left,
// as npq_ref_2d), then the reference's search inside subtree j.
template <int S>
__device__ __forceinline__ NpqD npq_ref_2d_sub(const MeshView &m, int root, double px, double py,
                                               int j, int *stk, int stride) {
    constexpr int D = S == 64 ? 6 : S == 32 ? 5 : S == 16 ? 4 : S == 8 ? 3 : S == 4 ? 2 : S == 2 ? 1 : 0;
    static_assert((1 << D) == S, "S must be a power of two up to 64");
    int curr = root;
    for (int l = 0; l < D; ++l) {
        const NodeD2 n = m.nd2[curr];
        if (n.count > 0) {
            if ((j & ((1 << (D - l)) - 1)) != 0) {
                NpqD r;
                r.d2 = INFINITY; r.px = r.py = r.pz = 0.0; r.s = r.t = 0.0; r.prim = -1; r.slot = -1;
                return r;
            }
            break;
        }
        const int left = curr + 1, right = n.idx;
        const bool left_near = box_d2_2(m.nd2[left], px, py) <= box_d2_2(m.nd2[right], px, py);
        const bool far = ((j >> (D - 1 - l)) & 1) != 0;
        curr = far == left_near ? right : left;
    }
    return npq_ref_2d(m, curr, px, py, stk, stride);
}

// Points whose start query needs more than kClassifyBudget nodes in some
// subtree of k_classify3 (points deep inside a closed surface, where nearly
// every primitive is about as far) are queued for k_classify3_deep, which
// searches each of them on a whole wavefront.
constexpr int kClassifyBudget = 256;
constexpr int kClassify2Split = 16;  // lanes a point in k_classify2s
struct DeepQueue {
    unsigned long long *count;
    int64_t *point;    // the point's index in the launch
    int32_t *rays;     // its first three casts (classify_tail's first_rays)
};

// 3D classification with its two searches spread over a block: 64 points
// per block, wavefront j < S runs subtree j of their start queries and
// wavefront S + k their ray cast k (k < 3; the inside stream's draws 2k and
// 2k + 1), so a start query that must visit much of the tree (points deep
// inside a closed surface) runs on S lanes, and no wavefront mixes searches.
// Wavefront 0 then merges the subtrees' results and finishes the points.
template <int S>
__global__ __launch_bounds__((S + 3) * 64) void k_classify3(MeshView m, EqView e, GridSpec g,
                                                          DomainView dom, const double *plist,
                                                          const int64_t *sidx, int64_t n,
                                                          uint64_t seed,
                                                          int max_ray_attempts, uint8_t *loc,
                                                          StartInfo *starts, int32_t *amb,
                                                          double *bvalue, DeepQueue dq) {
    extern __shared__ int lds_i[];
    __shared__ NpqD s_npq[S * 64];
    __shared__ int s_ray[3 * 64];
    int *stk = lds_i + threadIdx.x;
    const int stride = blockDim.x;
    const int lane = (int)(threadIdx.x & 63u), role = (int)(threadIdx.x >> 6);
    const int64_t i = (int64_t)blockIdx.x * 64 + lane;
    double x = 0.0, y = 0.0, z = 0.0;
    if (i < n) point_coords<3>(g, plist, i, x, y, z);
    const uint64_t ps = point_seed(seed, (sidx && i < n) ? (uint64_t)sidx[i] : (uint64_t)i);
    if (role < S) {
        NpqD q;
        q.d2 = INFINITY; q.px = q.py = q.pz = 0.0; q.s = q.t = 0.0; q.prim = -1; q.slot = -1;
        if (i < n) q = npq_ref_3d_sub<S>(m, dom.root64, x, y, z, role, stk, stride,
                                         dq.count ? kClassifyBudget : 0);
        s_npq[role * 64 + lane] = q;
    } else {
        int res = 0;
        if (i < n) {
            Xoshiro rng(splitmix64(ps ^ kInsideStream));
            for (int k = S; k < role; ++k) { rng.unit(); rng.unit(); }
            const double u = rng.unit(), v = rng.unit();
            const double zz = 2 * u - 1;
            const double phi = 2 * kPi * v;
            const double rr = sqrt(1 - zz * zz);
            bool ambiguous;
            const int hits = ray_cast_d(m, dom.root64, x, y, z, rr * cos(phi), rr * sin(phi), zz,
                                        ambiguous, stk, stride);
            res = ambiguous ? 2 : (hits & 1);
        }
        s_ray[(role - S) * 64 + lane] = res;
    }
    __syncthreads();
    if (role != 0 || i >= n) return;
    NpqD q = s_npq[lane];
    bool deep = q.slot == -2;
    for (int j = 1; j < S; ++j) {
        const NpqD o = s_npq[j * 64 + lane];
        deep = deep || o.slot == -2;
        if (o.d2 < q.d2) q = o;
    }
    int first_rays = 0;
    for (int k = 0; k < 3; ++k) {
        const int res = s_ray[k * 64 + lane];
        first_rays |= ((res & 1) << k) | (((res >> 1) & 1) << (3 + k));
    }
    if (deep) {
        const unsigned long long slot = atomicAdd(dq.count, 1ull);
        dq.point[slot] = i;
        dq.rays[slot] = first_rays;
        return;
    }
    classify_tail<3>(m, e, dom, i, x, y, z, q, ps, first_rays, max_ray_attempts, stk, stride,
                     loc, starts, amb, bvalue);
}

// The queued points of k_classify3, one wavefront each: lane j searches
// subtree j of 64 (the split of npq_ref_3d_sub), the least distance with the
// lowest j wins (the reference's result, as in k_classify3), and lane 0
// finishes the point with the casts k_classify3 made.  The grid has a
// wavefront for every point of the launch; those past the queue's length
// return at once.
__global__ __launch_bounds__(256) void k_classify3_deep(MeshView m, EqView e, GridSpec g,
                                                        DomainView dom, const double *plist,
                                                        const int64_t *sidx, uint64_t seed,
                                                        int max_ray_attempts, uint8_t *loc,
                                                        StartInfo *starts, int32_t *amb,
                                                        double *bvalue, DeepQueue dq) {
    extern __shared__ int lds_i[];
    int *stk = lds_i + threadIdx.x;
    const int stride = blockDim.x;
    const int lane = (int)(threadIdx.x & 63u);
    const unsigned long long w = (unsigned long long)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (w >= *dq.count) return;
    const int64_t i = dq.point[w];
    double x = 0.0, y = 0.0, z = 0.0;
    point_coords<3>(g, plist, i, x, y, z);
    NpqD q = npq_ref_3d_sub<64>(m, dom.root64, x, y, z, lane, stk, stride);
    // least distance, lowest subtree on ties
    int rank = lane;
    for (int off = 1; off < 64; off <<= 1) {
        NpqD o;
        o.d2 = __shfl_xor(q.d2, off); o.px = __shfl_xor(q.px, off); o.py = __shfl_xor(q.py, off);
        o.pz = __shfl_xor(q.pz, off); o.s = __shfl_xor(q.s, off); o.t = __shfl_xor(q.t, off);
        o.prim = __shfl_xor(q.prim, off); o.slot = __shfl_xor(q.slot, off);
        const int orank = __shfl_xor(rank, off);
        if (o.d2 < q.d2 || (o.d2 == q.d2 && orank < rank)) { q = o; rank = orank; }
    }
    if (lane != 0) return;
    const uint64_t ps = point_seed(seed, sidx ? (uint64_t)sidx[i] : (uint64_t)i);
    classify_tail<3>(m, e, dom, i, x, y, z, q, ps, dq.rays[w], max_ray_attempts, stk, stride,
                     loc, starts, amb, bvalue);
}

// 2D classification on S lanes a point: lane j searches subtree j of the
// start query (npq_ref_2d_sub) and sums the winding number's crossings of
// the segments s0 + j, s0 + j + S, ...; the least distance with the lowest
// j is the reference's search result (as in k_classify3_deep) and the
// integer crossing counts add up exactly to winding_number_d's.  Lane 0
// finishes the point.
template <int S>
__global__ __launch_bounds__(256) void k_classify2s(MeshView m, EqView e, GridSpec g,
                                                    DomainView dom, const double *plist,
                                                    const int64_t *sidx, int64_t n, uint64_t seed,
                                                    int max_ray_attempts, uint8_t *loc,
                                                    StartInfo *starts, int32_t *amb,
                                                    double *bvalue) {
    extern __shared__ int lds_i[];
    int *stk = lds_i + threadIdx.x;
    const int stride = blockDim.x;
    const int j = (int)(threadIdx.x % S);
    const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / S;
    if (i >= n) return;  // whole groups of S lanes
    double x = 0.0, y = 0.0, z = 0.0;
    point_coords<2>(g, plist, i, x, y, z);
    NpqD q = npq_ref_2d_sub<S>(m, dom.root64, x, y, j, stk, stride);
    int wn = 0;
    for (int sg = dom.s0 + j; sg < dom.s1; sg += S) {
        const SegD sd = m.segd[sg];
        const double area = (sd.s1x - sd.s0x) * (y - sd.s0y) - (x - sd.s0x) * (sd.s1y - sd.s0y);
        if (sd.s0y <= y) {
            if (sd.s1y > y && area > 0) wn++;
        } else {
            if (sd.s1y <= y && area < 0) wn--;
        }
    }
    int rank = j;
    for (int off = 1; off < S; off <<= 1) {
        NpqD o;
        o.d2 = __shfl_xor(q.d2, off); o.px = __shfl_xor(q.px, off); o.py = __shfl_xor(q.py, off);
        o.pz = __shfl_xor(q.pz, off); o.s = __shfl_xor(q.s, off); o.t = __shfl_xor(q.t, off);
        o.prim = __shfl_xor(q.prim, off); o.slot = __shfl_xor(q.slot, off);
        const int orank = __shfl_xor(rank, off);
        wn += __shfl_xor(wn, off);
        if (o.d2 < q.d2 || (o.d2 == q.d2 && orank < rank)) { q = o; rank = orank; }
    }
    if (j != 0) return;
    const uint64_t ps = point_seed(seed, sidx ? (uint64_t)sidx[i] : (uint64_t)i);
    classify_tail<2>(m, e, dom, i, x, y, z, q, ps, wn != 0 ? 1 : 0, max_ray_attempts, stk, stride,
                     loc, starts, amb, bvalue);
}

// --------------------------------------------------------------------------
// distance grid build: FP64 reference search from each cell centre (formed in
// FP32 exactly as the walker forms it), then the primitives within
// dc + 2 hd of the centre, counted (phase 0) and written (phase 1).
// --------------------------------------------------------------------------
__device__ __forceinline__ float round_down_f(double x) {
    float f = (float)x;
    if ((double)f > x) f = nextafterf(f, -INFINITY);
    return f;
}

// Primitives within sqrt(r2) of c: counted, and when `out` is given written
// as (slot, distance rounded down) in ascending distance (insertion sort).
template <int DIM>
__device__ int range_collect(const MeshView &m, int root, double cx, double cy, double cz,
                             double r2, int *stk, int stride, GridEntry *out, int cap) {
    int sp = 1;
    stk[0] = root;
    int count = 0;
    while (sp > 0) {
        const int curr = stk[(--sp) * stride];
        int cnt, idx;
        if constexpr (DIM == 3) {
            const NodeD3 n = m.nd3[curr];
            if (box_d2_3(n, cx, cy, cz) > r2) continue;
            cnt = n.count; idx = n.idx;
        } else {
            const NodeD2 n = m.nd2[curr];
            if (box_d2_2(n, cx, cy) > r2) continue;
            cnt = n.count; idx = n.idx;
        }
        if (cnt > 0) {
            for (int i = 0; i < cnt; ++i) {
                const int slot = idx + i;
                double nx, ny, nz, s, t, d;
                if constexpr (DIM == 3) d = eberly_d(m.trid[slot], cx, cy, cz, nx, ny, nz, s, t);
                else d = npq_seg_d(m.segd[slot], cx, cy, nx, ny, s);
                if (d <= r2) {
                    if (count >= cap) return cap + 1;
                    if (out) {
                        GridEntry e;
                        e.slot = slot;
                        e.dmin = round_down_f(sqrt(d));
                        int j = count;
                        while (j > 0 && out[j - 1].dmin > e.dmin) { out[j] = out[j - 1]; --j; }
                        out[j] = e;
                    }
                    ++count;
                }
            }
        } else {
            stk[(sp++) * stride] = idx;       // right
            stk[(sp++) * stride] = curr + 1;  // left
        }
    }
    return count;
}

template <int DIM>
__global__ __launch_bounds__(kBlock) void k_grid_build(MeshView m, GridInfo G, int root64,
                                                       double pad, int cap, double band,
                                                       int phase, GridCellF *cells,
                                                       int32_t *counts, double *rl2,
                                                       const int64_t *offsets, GridEntry *glist,
                                                       float4 *planes, double plane_tol,
                                                       unsigned int *n_planar) {
    extern __shared__ int lds_g[];
    int *stk = lds_g + threadIdx.x;
    const int n = G.nx * G.ny * G.nz;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int ix = i / (G.ny * G.nz), iy = (i / G.nz) % G.ny, iz = i % G.nz;
    const double cx = (double)(G.ox + ((float)ix + 0.5f) * G.h);
    const double cy = (double)(G.oy + ((float)iy + 0.5f) * G.h);
    const double cz = DIM == 3 ? (double)(G.oz + ((float)iz + 0.5f) * G.h) : 0.0;
    const int c = G.cell_off + i;
    if (phase == 0) {
        const NpqD q = DIM == 2 ? npq_ref_2d(m, root64, cx, cy, stk, blockDim.x)
                                : npq_ref_3d(m, root64, cx, cy, cz, stk, blockDim.x);
        const double dc = sqrt(q.d2);
        const double rl = (dc + pad) * (1.0 + 1e-6);
        // a cell beyond `band` is answered by its bound alone: no list
        const int cnt = dc < band ? range_collect<DIM>(m, root64, cx, cy, cz, rl * rl, stk,
                                                       blockDim.x, nullptr, cap)
                                  : cap + 1;
        cells[c].dcbits = __float_as_uint(round_down_f(dc)) & ~kGridLenMask;
        cells[c].first = 0u;
        cells[c].slot0 = -1;
        cells[c].dmin1 = INFINITY;
        counts[c] = cnt <= cap ? cnt : 0;
        rl2[c] = rl * rl;
    } else if (counts[c] > 0) {
        const int64_t off = offsets[c];
        const int cnt = range_collect<DIM>(m, root64, cx, cy, cz, rl2[c], stk, blockDim.x,
                                           glist + off, cap);
        cells[c].dcbits |= (uint32_t)cnt;
        cells[c].slot0 = glist[off].slot;
        cells[c].dmin1 = cnt > 1 ? glist[off + 1].dmin : INFINITY;
        // planar cell: two or more listed triangles, every vertex of each
        // within plane_tol of entry 0's plane (FP64).  (A one-triangle list
        // is answered exactly by its entry 0, from the lane's LDS copy.)
        bool planar = false;
        if (DIM == 3 && planes && cnt >= 2) {
            const TriD &T0 = m.trid[glist[off].slot];
            double nx = T0.e1y * T0.e2z - T0.e1z * T0.e2y;
            double ny = T0.e1z * T0.e2x - T0.e1x * T0.e2z;
            double nz = T0.e1x * T0.e2y - T0.e1y * T0.e2x;
            const double len = sqrt(nx * nx + ny * ny + nz * nz);
            planar = len > 0.0;
            if (planar) {
                nx /= len; ny /= len; nz /= len;
                const double d = nx * T0.v0x + ny * T0.v0y + nz * T0.v0z;
                for (int k = 0; k < cnt && planar; ++k) {
                    const TriD &T = m.trid[glist[off + k].slot];
                    const double h0 = nx * T.v0x + ny * T.v0y + nz * T.v0z - d;
                    const double h1 = h0 + nx * T.e1x + ny * T.e1y + nz * T.e1z;
                    const double h2 = h0 + nx * T.e2x + ny * T.e2y + nz * T.e2z;
                    planar = fabs(h0) <= plane_tol && fabs(h1) <= plane_tol && fabs(h2) <= plane_tol;
                }
                if (planar) {
                    planes[c] = make_float4((float)nx, (float)ny, (float)nz, (float)d);
                    atomicAdd(n_planar, 1u);
                }
            }
        }
        cells[c].first = (uint32_t)off | (planar ? kGridPlanar : 0u);
    }
}

// --------------------------------------------------------------------------
// closest boundary point of a point list: bvh_npq (include/wos/bvh.hpp:38-39)
// per lane.  FP64 runs the reference traversal; FP32 runs the walker's query
// (pruning factor cf) and projects onto the primitive it found in FP64.
// --------------------------------------------------------------------------
template <int DIM, int W>
__global__ __launch_bounds__(kBlock) void k_closest(MeshView m, int32_t root64, int32_t root32,
                                                    int domain, int fp32, float cf, const double *pts,
                                                    int64_t n, double *dist, double *lower,
                                                    double *nearest, int32_t *prim) {
    extern __shared__ __align__(16) unsigned char lds_c[];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = pts[i * DIM], y = pts[i * DIM + 1], z = DIM == 3 ? pts[i * DIM + 2] : 0.0;
    double nx = 0.0, ny = 0.0, nz = 0.0;
    if (!fp32) {
        int *stk = reinterpret_cast<int *>(lds_c) + threadIdx.x;
        const NpqD q = DIM == 2 ? npq_ref_2d(m, root64, x, y, stk, blockDim.x)
                                : npq_ref_3d(m, root64, x, y, z, stk, blockDim.x);
        dist[i] = sqrt(q.d2);
        lower[i] = dist[i];
        prim[i] = q.prim;
        nx = q.px; ny = q.py; nz = q.pz;
    } else {
        const StackRef stk{reinterpret_cast<StackEntry *>(lds_c) + threadIdx.x, (int)blockDim.x,
                           nullptr, 0, 1 << 30};
        const float fx = (float)x, fy = (float)y, fz = (float)z;
        QueryF q{};
        q.best_slot = -1;
        bool more = (query_begin<DIM>(q, m, root32, domain, cf, fx, fy, fz) & kQsMore) != 0u;
        while (more)
            more = (query_step<DIM, W>(q, m, fx, fy, fz, stk, nullptr, 0) & kQsMore) != 0u;
        dist[i] = sqrt((double)q.best);
        lower[i] = sqrt((double)fminf(q.best, q.lb));
        const int slot = q.best_slot;
        if (slot >= 0) {
            prim[i] = m.slot_prim[slot];
            double s, t;
            if (DIM == 2) npq_seg_d(m.segd[slot], x, y, nx, ny, s);
            else eberly_d(m.trid[slot], x, y, z, nx, ny, nz, s, t);
        } else {
            // the distance grid's bound ended the search: no primitive
            prim[i] = -1;
            nx = ny = nz = NAN;
        }
    }
    nearest[i * DIM] = nx;
    nearest[i * DIM + 1] = ny;
    if (DIM == 3) nearest[i * DIM + 2] = nz;
}

// --------------------------------------------------------------------------
// adjoint staging: one record per walk of a round (point major, like vals)
// --------------------------------------------------------------------------
struct AdjOut {
    double *scal;      // 6 per walk: d alpha, d src_value, d src_radius, d src_center[3]
    double *vw;        // 3 per walk: weight x blend coefficient of each primitive vertex
    int32_t *prim;     // terminal primitive per walk (-1: none)
};

// terminal contribution: d(weight g)/d alpha = weight g L; the boundary data
// enter linearly through the blend (1-s, s) / (1-s-t, s, t) of the exit primitive
__device__ __forceinline__ void store_adjoint(const AdjOut &a, int64_t item, WalkDer &D,
                                              double weight, double g, int prim, double s,
                                              double t) {
    D.d_alpha += weight * g * D.L;
    double *o = a.scal + (size_t)item * 6;
    o[0] = D.d_alpha; o[1] = D.d_q; o[2] = D.d_sigma;
    o[3] = D.d_c[0]; o[4] = D.d_c[1]; o[5] = D.d_c[2];
    double *v = a.vw + (size_t)item * 3;
    v[0] = weight * (1.0 - s - t);
    v[1] = weight * s;
    v[2] = weight * t;
    a.prim[item] = prim;
}

template <int DIM, typename T, int SMODE, typename RNG>
__device__ double source_term_adj(const EqView &e, T px, T py, T pz, T R, T weight, RNG &rng,
                                  WalkDer &D);

// --------------------------------------------------------------------------
// one round of walks: active point k runs walks [wbeg[k], wbeg[k] + count_k);
// item i of the round is walk (i - off[k]) of the point with
// off[k] <= i < off[k+1], and its result goes to staging slot i (point major).
// off == nullptr: every point runs cw walks starting at w0.
// --------------------------------------------------------------------------
constexpr int kItemBlockShift = 6;   // items per entry of RoundView::kblk: 64

struct RoundView {
    const StartInfo *starts;
    const StartF *sf;        // inside id -> FP32 start copy
    const int32_t *inside;   // inside id -> point index
    const int32_t *pid;      // active k -> inside id
    int pid_identity;        // pid[k] == k for every k
    const int32_t *wbeg;     // active k -> first walk index (off != nullptr)
    const int64_t *off;      // exclusive prefix of counts, n + 1 entries, or nullptr
    const int32_t *kblk;     // (off != nullptr) item block b -> the k holding item b * 64
    int n;
    int cw;
    int w0;
    // i / cw for item indices i < 2^32 as (t + ((i - t) >> s1)) >> s2 with
    // t = mulhi(cw_m, i) (division by an invariant integer, Granlund and
    // Montgomery 1994; set by set_walks_per_point)
    uint32_t cw_m;
    int cw_s1, cw_s2;
    int64_t total;
    // when set (graph replay), n and total are read from device memory: the
    // inside count is known only on the device
    const int32_t *n_dev;
    const int64_t *total_dev;
};

// The active point k of item i and the walk index it runs.
__device__ __forceinline__ int round_point(const RoundView &r, int64_t i, uint32_t &walk) {
    int k;
    if (!r.off) {
        const uint32_t n = (uint32_t)i;
        const uint32_t t = __umulhi(r.cw_m, n);
        const uint32_t q = (t + ((n - t) >> r.cw_s1)) >> r.cw_s2;
        k = (int)q;
        walk = (uint32_t)r.w0 + (n - q * (uint32_t)r.cw);
    } else {
        k = r.kblk[i >> kItemBlockShift];
        while (r.off[k + 1] <= i) ++k;
        walk = (uint32_t)(r.wbeg[k] + (int)(i - r.off[k]));
    }
    return k;
}

// kblk of an uneven round: for every block b of 64 items the k with
// off[k] <= b * 64 < off[k + 1].  off (n + 1 entries, off[0] = 0) is strictly
// increasing, since every point listed in a round has walks in it.
__global__ __launch_bounds__(kBlock) void k_round_kblk(const int64_t *off, int n, int64_t nb,
                                                       int32_t *kblk) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const int64_t item = b << kItemBlockShift;
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= item) lo = mid;
        else hi = mid - 1;
    }
    kblk[b] = lo;
}

__device__ __forceinline__ int round_inside_id(const RoundView &r, int k) {
    return r.pid_identity ? k : r.pid[k];
}

__device__ __forceinline__ const StartInfo &round_item(const RoundView &r, int64_t i, uint32_t &walk) {
    return r.starts[r.inside[round_inside_id(r, round_point(r, i, walk))]];
}

// cw walks per point: the multiplier and shifts of round_point's division
// (cw >= 1; item indices stay below 2^32, the staging limit of a round)
void set_walks_per_point(RoundView &r, int cw) {
    int l = 0;
    while (((int64_t)1 << l) < (int64_t)cw) ++l;
    r.cw = cw;
    r.cw_m = (uint32_t)((((uint64_t)1 << 32) * (((uint64_t)1 << l) - (uint64_t)cw)) / (uint64_t)cw + 1);
    r.cw_s1 = l < 1 ? l : 1;
    r.cw_s2 = l > 1 ? l - 1 : 0;
}

// --------------------------------------------------------------------------
// kernel 2a: FP64 plain walker, persistent lanes with resumable searches
// --------------------------------------------------------------------------
// The grid holds as many lanes as the GPU keeps resident; each holds one walk
// at a time.  An iteration: the lanes whose search has finished end their
// walk or take their step (the step's draws, source term, weight and move,
// then the start of the next search); free lanes take the next walks of
// their wavefront's pool (ranked by mbcnt; the pool takes kFp64Chunk items at
// a time with one atomicAdd on `next`); then the searches advance by one unit
// per pass (query_d_step) until min_ready lanes are ready or none searches.
// A walk's draws and arithmetic, and each search's decisions, are those of
// one lane per walk running the reference's loop, so a walk's value and
// steps do not depend on the schedule.
// At least 3 wavefronts per SIMD (at most 168 VGPRs) for the 3D forward
// walks; 2D walks with a source term and adjoint replays may use 256 rather
// than spill.
constexpr int kFp64Chunk = 64;  // items a wavefront of k_walk_fp64 takes at a time
// k_walk_fp64's ready lanes per step phase: its search units (an FP64 node
// or primitive test) are dearer than the FP32 walker's, so it steps with
// fewer lanes ready (walker ms at 8 / 16 / 24 / 40: icosphere 110 / 107 /
// 109 / 115, torus 72 / 74 / 77 / 83, plate 385 / 370 / 379 / 399)
constexpr int kFp64MinReady = 16;

template <int DIM, bool SRC, bool SCR, bool ADJ, int SX = 0>
__global__ __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu(ADJ || DIM == 2 ? 2 : 3, 8)))
void k_walk_fp64(MeshView m, EqView e, RoundView rv, double eps, int max_steps, double *vals,
                 int32_t *steps_out, AdjOut adj, unsigned long long *next, int min_ready) {
    extern __shared__ int lds_i[];
    int *stk = lds_i + threadIdx.x;
    const int stride = blockDim.x;
    const int lane = __lane_id();
    const double alpha = e.alpha;
    bool active = false;     // the lane holds a walk
    bool searching = false;  // its search is in progress
    // wave-uniform pool of items [pool_next, pool_end); exhausted once `next`
    // has passed the last item
    bool exhausted = false;
    uint64_t pool_next = 0, pool_end = 0;
    int64_t item = 0;
    // the walk held by the lane.  The start record is read field by field
    // where needed (a register copy of all of it would stay live).
    uint32_t w = 0;
    const StartInfo *S = nullptr;
    double px = 0.0, py = 0.0, pz = 0.0, R = 0.0;
    uint32_t k0 = 0, k1 = 0;
    int root = 0;
    QueryD q{};
    double sample = 0.0, weight = 1.0;
    int steps = 0;
    WalkDer D{};
    // the walk's end: its value (and adjoint record) to its staging slot
    auto finish = [&]() {
        const bool hit = R > eps;
        const double g = steps == 0 ? S->g
                                    : eq_boundary(e, m, q.r.px, q.r.py, q.r.pz, q.r.prim, q.r.s, q.r.t);
        vals[item] = sample + weight * g;
        steps_out[item] = steps | (hit ? (int32_t)0x80000000 : 0);
        if constexpr (ADJ) {
            const int prim = steps == 0 ? S->prim : q.r.prim;
            const double s = steps == 0 ? S->s : q.r.s, t = steps == 0 ? S->t : q.r.t;
            store_adjoint(adj, item, D, weight, g, prim, s, DIM == 3 ? t : 0.0);
        }
        active = false;
    };
    for (;;) {
        // ---- finished searches: end the walk or take a step
        bool stepping = false;
        if (active && !searching) {
            if (R > eps && steps < max_steps) stepping = true;
            else finish();
        }
        // ---- free lanes take the next walks
        const uint64_t free_mask = __ballot(!active);
        if (free_mask != 0 && !exhausted && pool_next >= pool_end) {
            unsigned long long base = 0;
            if (lane == 0) base = atomicAdd(next, (unsigned long long)kFp64Chunk);
            base = __shfl(base, 0);
            if (base >= (unsigned long long)rv.total) {
                exhausted = true;
            } else {
                pool_next = base;
                const uint64_t end = base + (uint64_t)kFp64Chunk;
                pool_end = end < (uint64_t)rv.total ? end : (uint64_t)rv.total;
            }
        }
        if (free_mask != 0 && pool_next < pool_end) {
            if (!active) {
                const uint32_t lo = (uint32_t)free_mask, hi = (uint32_t)(free_mask >> 32);
                const int rank = __builtin_amdgcn_mbcnt_hi(hi, __builtin_amdgcn_mbcnt_lo(lo, 0));
                item = (int64_t)(pool_next + (uint64_t)rank);
                if ((uint64_t)item < pool_end) {
                    S = &round_item(rv, item, w);
                    px = S->x; py = S->y; pz = S->z; R = S->r;
                    k0 = S->key0; k1 = S->key1;
                    root = S->root64;
                    sample = 0.0; weight = 1.0;
                    steps = 0;
                    if constexpr (ADJ) D = WalkDer{};
                    active = true;
                    if (R > eps && steps < max_steps) stepping = true;
                    else finish();
                }
            }
            const uint64_t nfree = (uint64_t)__popcll(free_mask);
            pool_next = pool_next + nfree < pool_end ? pool_next + nfree : pool_end;
        }
        // ---- one walk step (include/wos/wos.hpp:68-127), then the next search
        if (stepping) {
            // the step's counter stream: the source term draws from slot
            // DIM-1 on, then the direction takes slots 0 .. DIM-2 (block 0 is
            // formed once when the source's draws stayed in it)
            StepStreamD rng(w, (uint32_t)steps, k0, k1, DIM - 1);
            if (SRC && e.source_kind != WOS_SRC_ZERO) {
                if constexpr (ADJ)
                    sample += source_term_adj<DIM, double, 1>(e, px, py, pz, R, weight, rng, D);
                else
                    sample += source_term<DIM, double, StepStreamD, SX, SCR>(e, px, py, pz, R, weight, rng);
            }
            if (SCR) {
                if constexpr (ADJ && DIM == 3)
                    if (e.kind == WOS_EQ_SCREENED) D.L += dlog_w3_dalpha(alpha, R);
                weight *= step_weight<DIM, double>(e.kind, alpha, R);
            }
            rng.seek(0);
            const double du = rng.unit();
            const double dv = DIM == 3 ? rng.unit() : 0.0;
            if (DIM == 2) {
                const double theta = du * 2 * kPi;
                px = px + R * cos(theta);
                py = py + R * sin(theta);
            } else {
                const double u = du, v = dv;
                const double z = 2 * u - 1, phi = 2 * kPi * v;
                const double rr = sqrt(1 - z * z);
                px = px + R * rr * cos(phi);
                py = py + R * rr * sin(phi);
                pz = pz + R * z;
            }
            ++steps;
            query_d_begin(q, root, stk);
            searching = true;
        }
        if (__ballot(active) == 0) {
            if (exhausted) break;
            continue;
        }
        // ---- search units until enough lanes are ready to step
        for (;;) {
            if (searching) {
                searching = query_d_step<DIM>(q, m, px, py, pz, stk, stride);
                if (!searching) R = sqrt(q.r.d2);
            }
            if (__ballot(searching) == 0) break;
            const uint64_t ready = __ballot(!searching && (active || !exhausted));
            if (__popcll(ready) >= min_ready) break;
        }
    }
}

// --------------------------------------------------------------------------
// kernel 2c: FP64 walker on the reference's streams, a lane per point
// --------------------------------------------------------------------------
// The reference draws a point's walks one after another from per-point
// xoshiro256++ streams (include/wos/runner.hpp:208-221): 2D a walk stream and
// a source stream, 3D one stream for both.  Walk i's draws start where walk
// i-1 stopped, so one lane runs all of a point's walks of a round in order,
// with the streams' state kept per inside point across rounds (8 words:
// walk stream, source stream).  Steps, sources, weights and queries are those
// of k_walk_fp64 (include/wos/wos.hpp:54-138), drawing in the reference's
// order: source sample (any source kind but none, a zero source included),
// then the step direction.

// a stream as the samplers' RNG interface
struct XoRef {
    static constexpr bool radius_first = false;
    Xoshiro *x;
    __device__ double unit() { return x->unit(); }
    __device__ double unit_open() { return x->unit_open(); }
};

__global__ __launch_bounds__(kBlock) void k_ref_seed(const StartInfo *starts, const int32_t *inside,
                                                     int64_t n, int dim, uint64_t *state) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const StartInfo &S = starts[inside[j]];
    const Xoshiro w(((uint64_t)S.key1 << 32) | S.key0);
    const Xoshiro src(dim == 2 ? (((uint64_t)S.skey1 << 32) | S.skey0) : 0ULL);
    uint64_t *o = state + 8 * j;
    o[0] = w.s0; o[1] = w.s1; o[2] = w.s2; o[3] = w.s3;
    o[4] = src.s0; o[5] = src.s1; o[6] = src.s2; o[7] = src.s3;
}

// A lane holds one point of the round and runs its walks in order; the
// searches are resumable (query_d_step), and the walker loop is that of
// k_walk_fp64: lanes whose search has finished end their walk (then start
// the point's next walk, or save the point's streams and take the next
// point of their wavefront's pool) or take their step together, and search
// passes run until kFp64MinReady lanes are ready.
template <int DIM, bool SRC, bool SCR, int SX = 0>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2, 8)))
void k_walk_ref(MeshView m, EqView e, RoundView rv, double eps, int max_steps, double *vals,
                int32_t *steps_out, uint64_t *state, unsigned long long *next, int min_ready) {
    extern __shared__ int lds_i[];
    int *stk = lds_i + threadIdx.x;
    const int stride = blockDim.x;
    const int lane = __lane_id();
    const double alpha = e.alpha;
    bool active = false;     // the lane holds a point
    bool searching = false;  // its walk's search is in progress
    bool exhausted = false;  // wave-uniform: every point is taken
    uint64_t pool_next = 0, pool_end = 0;
    // the point: its streams, its walks [base, base + cnt) of the round
    uint64_t *st = nullptr;
    Xoshiro ws, ss;
    XoRef walk{&ws};
    XoRef source{DIM == 2 ? &ss : &ws};
    const StartInfo *S = nullptr;
    int64_t base = 0;
    int cnt = 0, i = 0, root = 0;
    // the walk
    double px = 0.0, py = 0.0, pz = 0.0, R = 0.0;
    QueryD q{};
    double sample = 0.0, weight = 1.0;
    int steps = 0;
    auto start_walk = [&]() {
        px = S->x; py = S->y; pz = S->z; R = S->r;
        sample = 0.0; weight = 1.0;
        steps = 0;
    };
    auto store_walk = [&]() {
        const bool hit = R > eps;
        const double g = steps == 0 ? S->g
                                    : eq_boundary(e, m, q.r.px, q.r.py, q.r.pz, q.r.prim, q.r.s, q.r.t);
        vals[base + i] = sample + weight * g;
        steps_out[base + i] = steps | (hit ? (int32_t)0x80000000 : 0);
    };
    auto save_streams = [&]() {
        st[0] = ws.s0; st[1] = ws.s1; st[2] = ws.s2; st[3] = ws.s3;
        st[4] = ss.s0; st[5] = ss.s1; st[6] = ss.s2; st[7] = ss.s3;
        active = false;
    };
    for (;;) {
        // ---- finished searches: end the walk or take a step
        bool stepping = false;
        if (active && !searching) {
            if (R > eps && steps < max_steps) {
                stepping = true;
            } else {
                store_walk();
                if (++i < cnt) {
                    start_walk();
                    stepping = true;  // the point's start needs steps (checked when taken)
                } else {
                    save_streams();
                }
            }
        }
        // ---- free lanes take the next points
        const uint64_t free_mask = __ballot(!active);
        if (free_mask != 0 && !exhausted && pool_next >= pool_end) {
            unsigned long long b0 = 0;
            if (lane == 0) b0 = atomicAdd(next, (unsigned long long)kFp64Chunk);
            b0 = __shfl(b0, 0);
            if (b0 >= (unsigned long long)rv.n) {
                exhausted = true;
            } else {
                pool_next = b0;
                const uint64_t end = b0 + (uint64_t)kFp64Chunk;
                pool_end = end < (uint64_t)rv.n ? end : (uint64_t)rv.n;
            }
        }
        if (free_mask != 0 && pool_next < pool_end) {
            if (!active) {
                const uint32_t lo = (uint32_t)free_mask, hi = (uint32_t)(free_mask >> 32);
                const int rank = __builtin_amdgcn_mbcnt_hi(hi, __builtin_amdgcn_mbcnt_lo(lo, 0));
                const uint64_t kk = pool_next + (uint64_t)rank;
                if (kk < pool_end) {
                    const int k = (int)kk;
                    const int j = round_inside_id(rv, k);
                    base = rv.off ? rv.off[k] : (int64_t)k * rv.cw;
                    cnt = rv.off ? (int)(rv.off[k + 1] - rv.off[k]) : rv.cw;
                    S = &rv.starts[rv.inside[j]];
                    root = S->root64;
                    st = state + 8 * (int64_t)j;
                    ws.s0 = st[0]; ws.s1 = st[1]; ws.s2 = st[2]; ws.s3 = st[3];
                    ss.s0 = st[4]; ss.s1 = st[5]; ss.s2 = st[6]; ss.s3 = st[7];
                    i = 0;
                    start_walk();
                    active = true;
                    if (R > eps && steps < max_steps) {
                        stepping = true;
                    } else {
                        // every walk of the point starts at its end: no draws
                        for (; i < cnt; ++i) store_walk();
                        save_streams();
                    }
                }
            }
            const uint64_t nfree = (uint64_t)__popcll(free_mask);
            pool_next = pool_next + nfree < pool_end ? pool_next + nfree : pool_end;
        }
        // ---- one walk step in the reference's draw order, then the search
        if (stepping) {
            if (SRC) sample += source_term<DIM, double, XoRef, SX, SCR>(e, px, py, pz, R, weight, source);
            if (SCR) weight *= step_weight<DIM, double>(e.kind, alpha, R);
            if (DIM == 2) {
                const double theta = walk.unit() * 2 * kPi;
                px = px + R * cos(theta);
                py = py + R * sin(theta);
            } else {
                const double u = walk.unit(), v = walk.unit();
                const double z = 2 * u - 1, phi = 2 * kPi * v;
                const double rr = sqrt(1 - z * z);
                px = px + R * rr * cos(phi);
                py = py + R * rr * sin(phi);
                pz = pz + R * z;
            }
            ++steps;
            query_d_begin(q, root, stk);
            searching = true;
        }
        if (__ballot(active) == 0) {
            if (exhausted) break;
            continue;
        }
        // ---- search units until enough lanes are ready to step
        for (;;) {
            if (searching) {
                searching = query_d_step<DIM>(q, m, px, py, pz, stk, stride);
                if (!searching) R = sqrt(q.r.d2);
            }
            if (__ballot(searching) == 0) break;
            const uint64_t ready = __ballot(!searching && (active || !exhausted));
            if (__popcll(ready) >= min_ready) break;
        }
    }
}

// FP32 source term.  Same draws as source_term (the step stream from slot
// DIM-1 on, `rng`), but the Green function takes the sample's distance from
// the centre in closed form (R sqrt(u) in 2D, R cbrt(w) in 3D) rather than
// from rounded coordinates, and a constant source skips the direction
// entirely: only the radius draw is made (the first source slot, which
// shares its Philox block with the step direction).
// SMODE: 1 uniform, 2 Green, 3 GREEN_MIS, 4 Green with WOS_SRC_FIELD,
// 5 uniform with WOS_SRC_FIELD (the field cases 2D only).
template <int DIM, int SMODE, bool SCR = true>
__device__ __forceinline__ double source_term_f(const EqView &e, float px, float py, float pz,
                                                float R, float weight, StepStreamF &rng,
                                                bool use_s3 = false, Scr3 s3 = Scr3{}) {
    constexpr bool FLD = SMODE == 5;
    if (SMODE == 2) return source_term<DIM, float, StepStreamF, 0, SCR>(e, px, py, pz, R, weight, rng);
    if constexpr (SMODE == 4)
        return source_term<DIM, float, StepStreamF, 2, SCR>(e, px, py, pz, R, weight, rng);
    if constexpr (SMODE == 3 && DIM == 2)
        return source_term_mis<float, StepStreamF, SCR>(e, px, py, R, weight, rng);
    if (DIM == 2 && !eq_source_may_intersect<float>(e, px, py, R)) return 0.0;
    const bool screened = SCR && e.kind == WOS_EQ_SCREENED;
    const float alpha = (float)e.alpha;
    float f, dist;
    if (e.source_kind == WOS_SRC_CONSTANT) {
        const float ro = rng.unit_open();
        dist = DIM == 2 ? R * Fx<float>::sqrt(ro) : R * Fx<float>::cbrt(ro);
        f = (float)e.src_value;
    } else {
        float sx, sy, sz = 0.0f;
        if (DIM == 2) {
            dist = R * Fx<float>::sqrt(rng.unit_open());
            const float theta = rng.unit();
            float sn, cs;
            __sincosf(theta * (float)(2.0 * kPi), &sn, &cs);
            sx = __builtin_fmaf(dist, cs, px);
            sy = __builtin_fmaf(dist, sn, py);
        } else {
            const float ww = rng.unit_open(), u = rng.unit(), v = rng.unit();
            const float z = 2.0f * u - 1.0f;
            const float surf = Fx<float>::sqrt(fmaxf(0.0f, 1.0f - z * z));
            dist = R * Fx<float>::cbrt(ww);
            float sn, cs;
            __sincosf(v * (float)(2.0 * kPi), &sn, &cs);
            sx = __builtin_fmaf(dist * surf, cs, px);
            sy = __builtin_fmaf(dist * surf, sn, py);
            sz = __builtin_fmaf(dist, z, pz);
        }
        f = eq_source<float, FLD>(e, sx, sy, sz, DIM);
    }
    if (f == 0.0f || !(dist > 0.0f)) return 0.0;
    float volume, G;
    if (DIM == 2) {
        volume = (float)kPi * R * R;
        G = screened ? screened_green_2d<float>(alpha, R, dist) : poisson_green_2d<float>(R, dist);
    } else {
        volume = (float)(4.0 * kPi / 3.0) * R * R * R;
        if (use_s3) G = scr3_green(s3, alpha, R, dist);
        else G = screened ? screened_green_3d<float>(alpha, R, dist) : poisson_green_3d<float>(R, dist);
    }
    return (double)(weight * volume * f * G);
}

// Source term plus its parameter derivatives (adjoint replay).  Same draws
// and same value as source_term (T = double) / source_term_f (T = float);
// the derivatives are formed in FP64.
template <int DIM, typename T, int SMODE, typename RNG>
__device__ double source_term_adj(const EqView &e, T px, T py, T pz, T R, T weight, RNG &rng,
                                  WalkDer &D) {
    constexpr bool F32 = sizeof(T) == 4;
    const bool screened = e.kind == WOS_EQ_SCREENED;
    const T alpha = (T)e.alpha;
    if (DIM == 2 && !eq_source_may_intersect<T>(e, px, py, R)) return 0.0;
    T sx = px, sy = py, sz = pz, dist = (T)0;
    if (SMODE == 2 || e.source_mode == WOS_SOURCE_GREEN) {
        T radius;
        if (DIM == 2) {
            radius = screened ? screened_radius_2d<T>(alpha, R, rng) : poisson_radius_2d<T>(R, rng);
            const T angle = (T)2 * (T)kPi * rng.unit();
            sx = px + radius * Fx<T>::cos(angle);
            sy = py + radius * Fx<T>::sin(angle);
        } else {
            radius = screened ? screened_radius_3d<T>(alpha, R, rng) : poisson_radius_3d<T>(R, rng);
            const T z = (T)2 * rng.unit() - (T)1;
            const T angle = (T)2 * (T)kPi * rng.unit();
            const T xy = Fx<T>::sqrt((T)1 - z * z);
            sx = px + radius * xy * Fx<T>::cos(angle);
            sy = py + radius * xy * Fx<T>::sin(angle);
            sz = pz + radius * z;
        }
        const T f = eq_source<T>(e, sx, sy, sz, DIM);
        if (f == (T)0) return 0.0;
        T mass;
        if (DIM == 2) mass = screened ? screened_mass_2d<T>(alpha, R) : R * R / (T)4;
        else mass = screened ? screened_mass_3d<T>(alpha, R) : R * R / (T)6;
        const double S = (double)(weight * mass * f);
        add_source_der<T>(D, e, DIM, (double)(weight * mass), S, sx, sy, sz, 0.0);
        return S;
    }
    // uniform ball point, radius first (the counter streams' order)
    if constexpr (F32) {
        if (e.source_kind == WOS_SRC_CONSTANT) {
            const float ro = rng.unit_open();
            dist = DIM == 2 ? R * Fx<float>::sqrt(ro) : R * Fx<float>::cbrt(ro);
        } else if (DIM == 2) {
            dist = R * Fx<float>::sqrt(rng.unit_open());
            const float theta = rng.unit();
            float sn, cs;
            __sincosf(theta * (float)(2.0 * kPi), &sn, &cs);
            sx = __builtin_fmaf(dist, cs, px);
            sy = __builtin_fmaf(dist, sn, py);
        } else {
            const float ww = rng.unit_open(), u = rng.unit(), v = rng.unit();
            const float z = 2.0f * u - 1.0f;
            const float surf = Fx<float>::sqrt(fmaxf(0.0f, 1.0f - z * z));
            dist = R * Fx<float>::cbrt(ww);
            float sn, cs;
            __sincosf(v * (float)(2.0 * kPi), &sn, &cs);
            sx = __builtin_fmaf(dist * surf, cs, px);
            sy = __builtin_fmaf(dist * surf, sn, py);
            sz = __builtin_fmaf(dist, z, pz);
        }
    } else {
        if (DIM == 2) {
            const T r = R * sqrt(rng.unit_open());
            const T theta = rng.unit() * (T)2 * (T)kPi;
            sx = px + r * cos(theta);
            sy = py + r * sin(theta);
            const T dx = px - sx, dy = py - sy;
            dist = sqrt(dx * dx + dy * dy);
        } else {
            const T ww = rng.unit_open(), u = rng.unit(), v = rng.unit();
            const T z = (T)2 * u - (T)1, phi = (T)2 * (T)kPi * v;
            const T surf = sqrt((T)1 - z * z), vol = R * cbrt(ww);
            sx = px + vol * surf * cos(phi);
            sy = py + vol * surf * sin(phi);
            sz = pz + vol * z;
            const T dx = px - sx, dy = py - sy, dz = pz - sz;
            dist = sqrt(dx * dx + dy * dy + dz * dz);
        }
    }
    const T f = e.source_kind == WOS_SRC_CONSTANT ? (T)e.src_value : eq_source<T>(e, sx, sy, sz, DIM);
    if (f == (T)0 || (F32 && !(dist > (T)0))) return 0.0;
    T volume, G;
    if (DIM == 2) {
        volume = (T)kPi * R * R;
        G = screened ? screened_green_2d<T>(alpha, R, dist) : poisson_green_2d<T>(R, dist);
    } else {
        volume = F32 ? (T)(4.0 * kPi / 3.0) * R * R * R : (T)4 * (T)kPi * (R * R * R) / (T)3;
        G = screened ? screened_green_3d<T>(alpha, R, dist) : poisson_green_3d<T>(R, dist);
    }
    const double S = (double)(weight * volume * f * G);
    double dlogG = 0.0;
    if (DIM == 3 && screened)
        dlogG = F32 ? (double)dlog_g3_dalpha_f((float)alpha, (float)R, (float)dist)
                    : dlog_g3_dalpha(e.alpha, (double)R, (double)dist);
    add_source_der<T>(D, e, DIM, (double)(weight * volume * G), S, sx, sy, sz, dlogG);
    return S;
}

// zero n 64-bit words (a kernel rather than a memset node, so that the
// launch sequence can be captured into a graph and replayed)
__global__ void k_zero_u64(unsigned long long *p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

// Launch fields that only the refill and the end-of-walk stores read.  The
// walker reads them from this record in device memory where it uses them
// (scalar loads through a pointer the compiler cannot hoist out of the
// loop), so they do not hold scalar registers across the loop; held as
// kernel arguments they were loaded once and spilled.  k_init_parts writes
// the record before each launch.
struct WalkCold {
    RoundView rv;
    unsigned long long *counter;
    double *vals;
    int32_t *steps_out;
    float4 *term_pw;
    int32_t *term_slot;
    int parts;
    int chunk;
};
typedef __attribute__((address_space(4))) const WalkCold ConstWalkCold;
static_assert(offsetof(WalkCold, rv) == 0, "cold_round reads the record's first bytes");
// the walker's counter buffer: one counter per part, 16 words apart, then
// the cold record
constexpr size_t kCounterWords = 16 * kMaxParts;
constexpr size_t kCounterBytes = kCounterWords * sizeof(unsigned long long) + sizeof(WalkCold);

// the cold record through a pointer opaque to the compiler (an empty asm
// that "changes" it): loads through it stay where they are written
__device__ __forceinline__ ConstWalkCold *cold_record(const WalkCold *p) {
    asm volatile("" : "+s"(p));
    return (ConstWalkCold *)p;
}
// the record's RoundView, word by word (scalar loads of the words used)
__device__ __forceinline__ RoundView cold_round(ConstWalkCold *K) {
    static_assert(sizeof(RoundView) % 4 == 0, "whole 32-bit words");
    typedef __attribute__((address_space(4))) const uint32_t ConstWord;
    const ConstWord *src = (const ConstWord *)K;
    uint32_t w[sizeof(RoundView) / 4];
#pragma unroll
    for (size_t k = 0; k < sizeof(RoundView) / 4; ++k) w[k] = src[k];
    RoundView rv;
    __builtin_memcpy(&rv, w, sizeof(rv));
    return rv;
}

// first item of every part of the walker's item range; thread 0 also
// writes the launch's cold record
__global__ void k_init_parts(unsigned long long *counter, int parts, int64_t total,
                             const int64_t *total_dev, WalkCold cold, WalkCold *cold_dst) {
    const int p = threadIdx.x;
    if (p == 0) *cold_dst = cold;
    if (p >= parts) return;
    const uint64_t t = total_dev ? (uint64_t)*total_dev : (uint64_t)total;
    counter[16 * p] = t * (uint64_t)p / (uint64_t)parts;
}


// Inside-point compaction on the device (graph replay, where the host does
// not see the classification): per block of kBlock points its number of
// inside points, one scan over the blocks, then every block writes its
// inside points in ascending order -- the same list the host builds.
__global__ __launch_bounds__(kBlock) void k_inside_count(const uint8_t *loc, int64_t n,
                                                         int32_t *bcount) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = __syncthreads_count(i < n && loc[i] == 1);
    if (threadIdx.x == 0) bcount[blockIdx.x] = c;
}

__global__ __launch_bounds__(1024) void k_inside_scan(const int32_t *bcount, int64_t nb,
                                                      int64_t *boff, int32_t *n_in,
                                                      int64_t *total, int64_t walks_per_point) {
    __shared__ int64_t part[1024];
    const int t = threadIdx.x;
    const int64_t per = (nb + 1023) / 1024;
    const int64_t b0 = min(nb, (int64_t)t * per), b1 = min(nb, b0 + per);
    int64_t sum = 0;
    for (int64_t b = b0; b < b1; ++b) sum += bcount[b];
    part[t] = sum;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        const int64_t v = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int64_t run = t > 0 ? part[t - 1] : 0;
    for (int64_t b = b0; b < b1; ++b) { boff[b] = run; run += bcount[b]; }
    if (t == 1023) {
        *n_in = (int32_t)part[1023];
        *total = part[1023] * walks_per_point;
    }
}

// FP32 start copies of the inside points (n_dev: inside count on the device)
__global__ __launch_bounds__(kBlock) void k_start_f(const StartInfo *starts, const int32_t *inside,
                                                    int64_t n, const int32_t *n_dev, StartF *sf) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= (n_dev ? (int64_t)*n_dev : n)) return;
    const StartInfo &S = starts[inside[j]];
    StartF f;
    f.x = (float)S.x;
    f.y = (float)S.y;
    f.z = (float)S.z;
    f.r = (float)S.r;
    f.slot = S.slot;
    f.domain = S.domain;
    f.key0 = S.key0;
    f.key1 = S.key1;
    sf[j] = f;
}

// The inside points in sweep order: inside[j] = index of the j-th inside
// point, and (sf != nullptr) its FP32 start copy sf[j] (k_start_f's record).
__global__ __launch_bounds__(kBlock) void k_inside_scatter(const uint8_t *loc, int64_t n,
                                                           const int64_t *boff, int32_t *inside,
                                                           const StartInfo *starts, StartF *sf) {
    __shared__ int scan[kBlock];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int f = (i < n && loc[i] == 1) ? 1 : 0;
    scan[threadIdx.x] = f;
    __syncthreads();
    for (int d = 1; d < kBlock; d <<= 1) {
        const int v = (int)threadIdx.x >= d ? scan[threadIdx.x - d] : 0;
        __syncthreads();
        scan[threadIdx.x] += v;
        __syncthreads();
    }
    if (!f) return;
    const int64_t j = boff[blockIdx.x] + scan[threadIdx.x] - 1;
    inside[j] = (int32_t)i;
    if (sf) {
        const StartInfo &S = starts[i];
        StartF r;
        r.x = (float)S.x;
        r.y = (float)S.y;
        r.z = (float)S.z;
        r.r = (float)S.r;
        r.slot = S.slot;
        r.domain = S.domain;
        r.key0 = S.key0;
        r.key1 = S.key1;
        sf[j] = r;
    }
}

// the per-point statistics of a solve and its failed-walk counter to zero,
// in one launch (a kernel, not memset nodes, so that it can be captured)
__global__ __launch_bounds__(kBlock) void k_zero_stats(double *mean, double *m2, double *count,
                                                       double *ssum, int32_t *hits,
                                                       unsigned long long *failed, int64_t n) {
    const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i0 == 0) *failed = 0ull;
    for (int64_t i = i0; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        mean[i] = 0.0;
        m2[i] = 0.0;
        count[i] = 0.0;
        ssum[i] = 0.0;
        hits[i] = 0;
    }
}

// --------------------------------------------------------------------------
// kernel 2b: FP32 persistent walker
// --------------------------------------------------------------------------
// Each lane holds one walk at a time and advances its closest-point query by
// one unit (node or primitive) per trip of the inner loop.  When min_ready
// lanes of the wave have finished their queries, or no lane is searching any
// more, the finished lanes take their walk step together: end the walk or go
// on, refill from the wave's pool of work items (ballot + mbcnt rank, one
// atomicAdd per chunk), move to the next sphere point and begin the next
// query.  Step queries prune with the factor cf; a walk that the relaxed
// bound would end, or that reaches max_steps, searches again exactly first,
// so the stopping rule and the terminal projection see the exact nearest
// primitive.
struct PersistArgs {
    RoundView rv;
    const WalkCold *cold;  // rv, counter, parts, chunk, vals, steps_out, term_*
    int max_steps;
    double eps;
    float eps_f;
    float shrink_rel;
    float cf;            // pruning factor of step queries, (1 - relax)^2
    int min_ready;       // finished lanes that start a walk-step phase
    int refill_min;      // free lanes that start a refill (all free: always)
    unsigned long long *counter;  // one per part, 16 words apart (own cache line)
    int parts;           // item range split into this many contiguous parts
    int chunk;
    int n_cached;        // internal BVH nodes staged in LDS (breadth-first prefix)
    int cache_bytes;     // LDS bytes reserved for them (16 B aligned)
    int tri_off;         // 3D: LDS offset of the lanes' triangle copies (-1: none)
    StackEntry *gstack;  // traversal stacks in global memory: entry k of lane g at
    int gstride;         // gstack[k * gstride + g], gstride = the launch's lanes
    int klds;            // stack entries per lane held in LDS (the first ones)
    int stack_off;       // LDS offset of those entries
    double *vals;
    int32_t *steps_out;
    unsigned long long *units;  // 7 query-unit counters, or nullptr
    // forward walks with g != 0: the terminal projection is left to the
    // merge (terminal_value), which reads the end point and weight (term_pw)
    // and the nearest primitive's slot (term_slot, -1: value final) of
    // every walk
    float4 *term_pw;
    int32_t *term_slot;
};

// Forward walks: at least 5 wavefronts per SIMD (at most 96 VGPRs) in 3D
// without Green-function sampling, which fit without spilling, else 4 (at
// most 128 VGPRs; the 3D Green samplers would spill at 96).  Adjoint
// replays carry the FP64 parameter derivatives and may use up to 256 VGPRs
// rather than spill them.
// PL: the grid has planar cells (3D), whose relaxed queries end at the
// plane distance (query_begin).
template <int DIM, int SMODE, bool SCR, bool ADJ, int W, bool PL>
__global__ __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu(ADJ ? 2 : (DIM == 3 && SMODE != 2 ? 5 : 4), 8)))
void k_walk_fp32(MeshView m, EqView e, PersistArgs a, AdjOut adj) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    // stage the top of the tree once per block
    {
        const int4 *src = W == 4 ? (DIM == 3 ? reinterpret_cast<const int4 *>(m.nw3)
                                             : reinterpret_cast<const int4 *>(m.nw2))
                                 : (DIM == 3 ? reinterpret_cast<const int4 *>(m.nf3)
                                             : reinterpret_cast<const int4 *>(m.nf2));
        int4 *dst = reinterpret_cast<int4 *>(lds_raw);
        constexpr int kNodeBytes = W == 4 ? (DIM == 3 ? (int)sizeof(NodeF3W) : (int)sizeof(NodeF2W))
                                          : (DIM == 3 ? (int)sizeof(NodeF3) : (int)sizeof(NodeF2));
        const int n16 = a.n_cached * kNodeBytes / 16;
        for (int i = threadIdx.x; i < n16; i += blockDim.x) dst[i] = src[i];
        __syncthreads();
    }
    const void *cache = lds_raw;
    // the BVH traversal stack (needed only where the distance grid cannot
    // answer, rare with a grid) lives in global memory, so that the LDS does
    // not limit occupancy for deep trees
    const StackRef stk{reinterpret_cast<StackEntry *>(lds_raw + a.stack_off) + threadIdx.x,
                       (int)blockDim.x, a.gstack + (size_t)blockIdx.x * blockDim.x + threadIdx.x,
                       a.gstride, a.klds};
    const TriCache tcp{DIM == 3 && a.tri_off > 0
                           ? reinterpret_cast<float4 *>(lds_raw + a.tri_off) + threadIdx.x
                           : nullptr,
                       (int)blockDim.x};
    const int lane = __lane_id();

    // lane state
    bool active = false;     // the lane holds a walk
    bool searching = false;  // its query is in progress
    float px = 0, py = 0, pz = 0, R = 0, weight = 1.0f, margin = 0.0f;
    double sample = 0.0;
    int steps = 0;
    int32_t root = 0, dom = 0;
    uint32_t w = 0, k0 = 0, k1 = 0;
    size_t out_index = 0;
    QueryF q{};
    q.best_slot = -1;
    q.tslot = -1;
    const float alpha = (float)e.alpha;
    WalkDer D{};
    // forward walks: the lane's walk has ended and its results are not yet
    // stored (g = 0: the value is the source sum; else the merge adds
    // weight x g at the FP64 nearest point, one lane per walk)
    bool pending = false;
    auto store_walk = [&]() {
        ConstWalkCold *K = cold_record(a.cold);
        const bool hit = R > a.eps_f;
        K->steps_out[out_index] = steps | (hit ? (int32_t)0x80000000 : 0);
        K->vals[out_index] = sample;
        int32_t *const term_slot = K->term_slot;
        if (term_slot) {
            K->term_pw[out_index] = make_float4(px, py, pz, weight);
            term_slot[out_index] = q.best_slot;
        }
        pending = false;
    };

    // wave-uniform pool of work items.  The item range is split into
    // `parts` contiguous parts (points are in sweep order, so a part is a
    // region of the domain); the hardware places workgroups on the XCDs
    // round-robin, so taking part blockIdx.x % parts first keeps each XCD's
    // L2 on one region.  An exhausted part sends the wave to the next one.
    uint64_t pool_next = 0, pool_end = 0;
    bool exhausted = false;
    uint64_t total;
    int part, parts_left;
    {
        ConstWalkCold *K = cold_record(a.cold);
        const int64_t *total_dev = K->rv.total_dev;
        total = total_dev ? (uint64_t)*total_dev : (uint64_t)K->rv.total;
        parts_left = K->parts;
        part = (int)(blockIdx.x % (unsigned)parts_left);
    }
    // wave-uniform query-unit counts (a.units != nullptr): queries, grid
    // bounds, list scans, list entries read, nodes, leaf primitives, list
    // primitives tested
    uint64_t u_q = 0, u_gb = 0, u_ls = 0, u_lent = 0, u_nd = 0, u_pr = 0, u_lprim = 0;

    for (;;) {
        bool began = false;
        // a walk that the relaxed bound would end searches again exactly
        // (cf = 1) and takes no step in this iteration
        bool requery = false;
        uint32_t qb = 0u;
        // ---- finished queries: radius, then end of walk or exact re-query
        if (active && !searching) {
            const bool exact = !(q.lb < q.best);
            R = Fx<float>::sqrt(exact ? q.best : q.lb) * a.shrink_rel - margin;
            const bool stop = !(R > a.eps_f) || steps >= a.max_steps;
            if (stop && !exact) {
                requery = true;
            } else if (stop) {
                if constexpr (ADJ) {
                    ConstWalkCold *K = cold_record(a.cold);
                    const bool hit = R > a.eps_f;
                    K->steps_out[out_index] = steps | (hit ? (int32_t)0x80000000 : 0);
                    // terminal value in FP64: nearest point on the nearest primitive
                    double g = 0.0, s = 0.0, t = 0.0;
                    int prim = -1;
                    const int slot = q.best_slot;
                    if (slot >= 0) {
                        prim = m.slot_prim[slot];
                        double nx, ny, nz = 0.0;
                        if (DIM == 2) npq_seg_d(m.segd[slot], (double)px, (double)py, nx, ny, s);
                        else eberly_d(m.trid[slot], (double)px, (double)py, (double)pz, nx, ny, nz, s, t);
                        g = eq_boundary(e, m, nx, ny, nz, prim, s, t);
                    }
                    K->vals[out_index] = sample + (double)weight * g;
                    store_adjoint(adj, (int64_t)out_index, D, (double)weight, g, prim, s, t);
                } else {
                    // stored when the lane is refilled (or the pool is empty):
                    // a few lanes end per iteration, and storing them there
                    // runs the stores about a third as often
                    pending = true;
                }
                active = false;
            }
        }

        // ---- refill free lanes from the wave pool
        const uint64_t free_mask = __ballot(!active);
        if (free_mask != 0 && !exhausted &&
            (__popcll(free_mask) >= a.refill_min || free_mask == ~0ull)) {
            ConstWalkCold *K = cold_record(a.cold);
            const int parts = K->parts;
            while (pool_next >= pool_end) {
                const uint64_t part_end = total * (uint64_t)(part + 1) / (uint64_t)parts;
                const int chunk = K->chunk;
                unsigned long long base = 0;
                if (lane == 0)
                    base = atomicAdd(K->counter + 16 * part, (unsigned long long)chunk);
                base = __shfl(base, 0);
                if (base < part_end) {
                    pool_next = base;
                    pool_end = base + (uint64_t)chunk;
                    if (pool_end > part_end) pool_end = part_end;
                } else if (--parts_left == 0) {
                    exhausted = true;
                    break;
                } else {
                    part = part + 1 == parts ? 0 : part + 1;
                }
            }
            if (!exhausted) {
                const uint32_t lo = (uint32_t)free_mask, hi = (uint32_t)(free_mask >> 32);
                const int rank = __builtin_amdgcn_mbcnt_hi(hi, __builtin_amdgcn_mbcnt_lo(lo, 0));
                const uint64_t item = pool_next + (uint64_t)rank;
                if (pending && item < pool_end) store_walk();
                if (!active && item < pool_end) {
                    const RoundView rv = cold_round(K);
                    const int j = round_inside_id(rv, round_point(rv, (int64_t)item, w));
                    const StartF S = rv.sf[j];
                    dom = S.domain;
                    const DomainF DF = domain_record(m.doms, dom);
                    px = S.x; py = S.y; pz = S.z;
                    margin = DF.shrink_abs;
                    R = S.r * a.shrink_rel - margin;
                    weight = 1.0f;
                    sample = 0.0;
                    steps = 0;
                    q.best_slot = S.slot;
                    root = DF.root32;
                    k0 = S.key0;
                    k1 = S.key1;
                    out_index = (size_t)item;
                    active = true;
                    if constexpr (ADJ) D = WalkDer{};
                    if (!(R > a.eps_f)) {
                        // zero-step walk: g at the start's nearest point
                        const StartInfo &F = rv.starts[rv.inside[j]];
                        K->vals[out_index] = F.g;
                        K->steps_out[out_index] = 0;
                        int32_t *const term_slot = K->term_slot;
                        if (!ADJ && term_slot) term_slot[out_index] = -1;
                        if constexpr (ADJ)
                            store_adjoint(adj, (int64_t)item, D, 1.0, F.g, F.prim, F.s,
                                          DIM == 3 ? F.t : 0.0);
                        active = false;
                    }
                }
                const uint64_t nfree = (uint64_t)__popcll(free_mask);
                pool_next = pool_next + nfree < pool_end ? pool_next + nfree : pool_end;
            }
        }
        if (exhausted && pending) store_walk();

        // ---- one walk step (include/wos/wos.hpp:68-127), then the next query.
        // A lane re-querying exactly does not step: its radius is the
        // relaxed one, and the exact result (which query_begin may already
        // complete) decides the walk's end next iteration.  Both kinds of
        // lane then begin their query in one call.
        const bool stepping = active && !searching && !requery;
        if (stepping) {
            // block 0 of the step's counter stream: the direction's slots and
            // the first source slots (a constant source needs no other block)
            const Words4 b0 = philox4(w, (uint32_t)steps, 0, 0, k0, k1);
            // 3D screened forward walks: the weight's and the Green
            // function's exponentials of alpha R formed once
            constexpr bool kScr3 = SCR && DIM == 3 && !ADJ;
            const bool scr3 = kScr3 && e.kind == WOS_EQ_SCREENED;
            Scr3 s3{0.0f, 0.0f, 0.0f};
            if (scr3) s3 = scr3_terms(alpha, R);
            if (SMODE != 0 && e.source_kind != WOS_SRC_ZERO) {
                StepStreamF src(w, (uint32_t)steps, k0, k1, DIM - 1, b0);
                if constexpr (ADJ)
                    sample += source_term_adj<DIM, float, SMODE>(e, px, py, pz, R, weight, src, D);
                else
                    sample += source_term_f<DIM, SMODE, SCR>(e, px, py, pz, R, weight, src, scr3, s3);
            }
            if (SCR) {
                if constexpr (ADJ && DIM == 3)
                    if (e.kind == WOS_EQ_SCREENED) D.L += (double)dlog_w3_dalpha_f(alpha, R);
                weight *= scr3 ? scr3_weight(s3) : step_weight<DIM, float>(e.kind, alpha, R);
            }
            if (DIM == 2) {
                const float u = u24(b0.w0);
                float sn, cs;
                __sincosf(u * (float)(2.0 * kPi), &sn, &cs);
                px = __builtin_fmaf(R, cs, px);
                py = __builtin_fmaf(R, sn, py);
            } else {
                const float u = u24(b0.w0), v = u24(b0.w1);
                const float z = 2.0f * u - 1.0f;
                const float rr = Fx<float>::sqrt(fmaxf(0.0f, 1.0f - z * z));
                float sn, cs;
                __sincosf(v * (float)(2.0 * kPi), &sn, &cs);
                px = __builtin_fmaf(R * rr, cs, px);
                py = __builtin_fmaf(R * rr, sn, py);
                pz = __builtin_fmaf(R, z, pz);
            }
            ++steps;
        }
        if (stepping || requery) {
            qb = query_begin<DIM, PL>(q, m, root, dom, requery ? 1.0f : a.cf, px, py, pz, tcp);
            searching = (qb & kQsMore) != 0u;
            began = true;
        }
        if (a.units) {
            u_q += __popcll(__ballot(began));
            u_gb += __popcll(__ballot(began && !q.indirect && !searching));
            u_ls += __popcll(__ballot(began && q.indirect));
            u_lprim += __popcll(__ballot(((qb >> kQsPrimShift) & 3u) != 0u));
        }
        if (__ballot(active) == 0) {
            if (exhausted) break;
            continue;
        }

        // ---- advance the queries until enough lanes are ready to step
        for (;;) {
            uint32_t did = 0u;
            if (searching) {
                did = query_step<DIM, W>(q, m, px, py, pz, stk, cache, a.n_cached);
                searching = (did & kQsMore) != 0u;
            }
            if (a.units) {
                const uint32_t ne = (did >> kQsEntryShift) & 3u, np = (did >> kQsPrimShift) & 3u;
                u_nd += __popcll(__ballot((did & kQsNode) != 0u));
                u_pr += __popcll(__ballot((did & kQsLeaf) != 0u));
                u_lent += __popcll(__ballot(ne & 1u)) + 2 * __popcll(__ballot(ne & 2u));
                u_lprim += __popcll(__ballot(np & 1u)) + 2 * __popcll(__ballot(np & 2u));
            }
            if (__ballot(searching) == 0) break;
            const uint64_t ready = __ballot(!searching && (active || !exhausted));
            if (__popcll(ready) >= a.min_ready) break;
        }
    }
    if (a.units && lane == 0) {
        atomicAdd(a.units + 0, (unsigned long long)u_q);
        atomicAdd(a.units + 1, (unsigned long long)u_gb);
        atomicAdd(a.units + 2, (unsigned long long)u_ls);
        atomicAdd(a.units + 3, (unsigned long long)u_lent);
        atomicAdd(a.units + 4, (unsigned long long)u_nd);
        atomicAdd(a.units + 5, (unsigned long long)u_pr);
        atomicAdd(a.units + 6, (unsigned long long)u_lprim);
    }
}

// --------------------------------------------------------------------------
// terminal value of an FP32 forward walk when g != 0: weight x g at the FP64
// nearest point of the walk's nearest primitive `slot` (from the FP32 end
// point; walks whose value is already final carry slot -1 and are skipped by
// the caller).  Formed in the FP32 merge, one lane per walk, where the walk's
// value is read anyway (in the persistent walker only the few ending lanes
// of a wavefront would run this FP64 code).
// --------------------------------------------------------------------------
template <int DIM>
__device__ __forceinline__ double terminal_value(const MeshView &m, const EqView &e, float4 p,
                                                 int slot) {
    const int prim = m.slot_prim[slot];
    double nx, ny, nz = 0.0, s = 0.0, t = 0.0;
    if (DIM == 2) npq_seg_d(m.segd[slot], (double)p.x, (double)p.y, nx, ny, s);
    else eberly_d(m.trid[slot], (double)p.x, (double)p.y, (double)p.z, nx, ny, nz, s, t);
    const double g = eq_boundary(e, m, nx, ny, nz, prim, s, t);
    return (double)p.w * g;
}

// --------------------------------------------------------------------------
// kernel 3: sequential Welford per point (include/wos/wos.hpp:169-180)
// --------------------------------------------------------------------------
// Walks whose value is NaN (a rejection sampler that gave up) are counted
// into *failed; they still enter the statistics, so the point's mean is NaN.
__global__ __launch_bounds__(kBlock) void k_welford(RoundView rv, const double *vals,
                                                    const int32_t *steps, double *mean, double *m2,
                                                    double *count, double *steps_sum,
                                                    int32_t *hits, unsigned long long *failed) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (rv.n_dev ? *rv.n_dev : rv.n)) return;
    const int p = rv.pid[k];
    const int64_t base = rv.off ? rv.off[k] : (int64_t)k * rv.cw;
    const int cnt = rv.off ? (int)(rv.off[k + 1] - rv.off[k]) : rv.cw;
    double mu = mean[p], M2 = m2[p], n = count[p], ts = steps_sum[p];
    int h = hits[p];
    unsigned nan_walks = 0u;
    // the point's walks in batches of kBatch: all loads of a batch are in
    // flight together, then the updates run in walk order (one lane per
    // point leaves too few wavefronts to hide one load latency per walk)
    constexpr int kBatch = 16;
    for (int j0 = 0; j0 < cnt; j0 += kBatch) {
        const int nb = cnt - j0 < kBatch ? cnt - j0 : kBatch;
        double xs[kBatch];
        int32_t ss[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            if (u < nb) {
                const size_t o = (size_t)(base + j0 + u);
                xs[u] = vals[o];
                ss[u] = steps[o];
            }
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            if (u < nb) {
                const double x = xs[u];
                const int32_t s = ss[u];
                nan_walks += x != x ? 1u : 0u;
                n = n + 1.0;
                const double delta_before = x - mu;
                mu += delta_before / n;
                const double delta_after = x - mu;
                M2 += delta_before * delta_after;
                ts += (double)(s & 0x7fffffff);
                h += (s < 0) ? 1 : 0;
            }
        }
    }
    mean[p] = mu; m2[p] = M2; count[p] = n; steps_sum[p] = ts; hits[p] = h;
    if (nan_walks) atomicAdd(failed, (unsigned long long)nan_walks);
}

// --------------------------------------------------------------------------
// kernel 3b: the FP32 walker's merge, one wavefront per point.  Lane l owns
// the point's walks l, l + 64, l + 128, ... and runs the same Welford update
// over them in walk order; the 64 lane states (count, mean, M2) are combined
// pairwise by Chan, Golub and LeVeque's formula (1979) in a fixed butterfly
// order (lane distance 32, 16, ..., 1; each pair combines its lower lane
// first, so both partners hold the same state).  Where a point's walks come
// in several rounds, the lane states persist between them (lane_mu, lane_m2:
// 64 per point; a lane's count follows from the point's), so every split of
// the walks into rounds, and every batch, gives the same bits; they equal
// the sequential merge to rounding (the FP64 walkers keep the sequential
// order, which defines their bits).  A point's serial chain is 1/64 as
// long: the kernel no longer waits on one FP64 division per walk.  TERM:
// forward walks ending where g is not zero add their terminal value
// (terminal_value) as they are merged.
// --------------------------------------------------------------------------
struct Moments { double n, mu, M2; };
// (n, mu, M2) of the union of two disjoint samples; a.n = 0 returns b exactly
__device__ __forceinline__ Moments chan_combine(const Moments &a, const Moments &b) {
    const double n = a.n + b.n;
    if (n == 0.0) return a;
    const double delta = b.mu - a.mu;
    const double fb = b.n / n;
    return Moments{n, a.mu + delta * fb, a.M2 + b.M2 + delta * delta * a.n * fb};
}

template <int DIM, bool TERM>
__global__ __launch_bounds__(kBlock) void k_merge_wave(RoundView rv, const double *vals,
                                                       const int32_t *steps, double *mean,
                                                       double *m2, double *count,
                                                       double *steps_sum, int32_t *hits,
                                                       unsigned long long *failed,
                                                       double *lane_mu, double *lane_m2,
                                                       MeshView m, EqView e,
                                                       const float4 *term_pw,
                                                       const int32_t *term_slot) {
    const int k = (int)(blockIdx.x * (kBlock / 64) + threadIdx.x / 64);  // wave-uniform
    const int lane = (int)(threadIdx.x & 63);
    if (k >= (rv.n_dev ? *rv.n_dev : rv.n)) return;
    const int p = rv.pid[k];
    const int64_t base = rv.off ? rv.off[k] : (int64_t)k * rv.cw;
    const int cnt = rv.off ? (int)(rv.off[k + 1] - rv.off[k]) : rv.cw;
    const int wb = rv.off ? rv.wbeg[k] : rv.w0;  // walk index of the round's first walk
    // the lane's state after the earlier rounds (walks 0 .. wb - 1)
    const int64_t nl = wb > lane ? (int64_t)(wb - lane + 63) / 64 : 0;
    const size_t ls = (size_t)p * 64 + (size_t)lane;
    Moments s{(double)nl, 0.0, 0.0};
    if (nl > 0 && lane_mu) { s.mu = lane_mu[ls]; s.M2 = lane_m2[ls]; }
    double ts = 0.0;
    int h = 0;
    unsigned nan_walks = 0u;
    for (int j = (lane - wb) & 63; j < cnt; j += 64) {
        double x = vals[base + j];
        const int32_t st = steps[base + j];
        if constexpr (TERM) {
            const int slot = term_slot[base + j];
            if (slot >= 0) x = x + terminal_value<DIM>(m, e, term_pw[base + j], slot);
        }
        nan_walks += x != x ? 1u : 0u;
        s.n = s.n + 1.0;
        const double delta_before = x - s.mu;
        s.mu += delta_before / s.n;
        s.M2 += delta_before * (x - s.mu);
        ts += (double)(st & 0x7fffffff);
        h += st < 0 ? 1 : 0;
    }
    if (lane_mu) { lane_mu[ls] = s.mu; lane_m2[ls] = s.M2; }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const Moments t{__shfl_xor(s.n, o), __shfl_xor(s.mu, o), __shfl_xor(s.M2, o)};
        const bool lower = (lane & o) == 0;
        s = lower ? chan_combine(s, t) : chan_combine(t, s);
        // integer-valued sums: exact in any order
        ts += __shfl_xor(ts, o);
        h += __shfl_xor(h, o);
        nan_walks += __shfl_xor(nan_walks, o);
    }
    if (lane == 0) {
        mean[p] = s.mu; m2[p] = s.M2; count[p] = s.n;
        steps_sum[p] += ts;
        hits[p] += h;
        if (nan_walks) atomicAdd(failed, (unsigned long long)nan_walks);
    }
}

// Merge of a round's walks: one wavefront per point for the FP32 walker,
// the sequential merge for the FP64 walkers.  lane_mu / lane_m2: the FP32
// merge's lane states (64 per point), needed when a point's walks may come
// in more than one round, else nullptr.  term_slot (FP32 forward walks with
// g != 0): the merge adds each walk's terminal value.
hipError_t launch_merge(hipStream_t stream, int precision, const RoundView &rv, int64_t n_max,
                        double *vals, int32_t *steps, double *mean, double *m2, double *count,
                        double *ssum, int32_t *hits, unsigned long long *failed,
                        double *lane_mu, double *lane_m2, int dim, const MeshView &m,
                        const EqView &e, const float4 *term_pw, const int32_t *term_slot) {
    if (precision == WOS_FP32_PERSISTENT) {
        constexpr int kWaves = kBlock / 64;
        const dim3 grid((unsigned)((n_max + kWaves - 1) / kWaves));
#define WOS_MERGE(D, T)                                                                     \
    hipLaunchKernelGGL((k_merge_wave<D, T>), grid, dim3(kBlock), 0, stream, rv, vals, steps, \
                       mean, m2, count, ssum, hits, failed, lane_mu, lane_m2, m, e, term_pw, \
                       term_slot)
        if (term_slot && dim == 2) WOS_MERGE(2, true);
        else if (term_slot) WOS_MERGE(3, true);
        else WOS_MERGE(3, false);
#undef WOS_MERGE
    } else {
        // one wavefront per block: a point's merge is a serial chain, so
        // the kernel wants its few wavefronts on as many CUs as possible
        hipLaunchKernelGGL(k_welford, dim3((unsigned)((n_max + 63) / 64)), dim3(64), 0, stream,
                           rv, vals, steps, mean, m2, count, ssum, hits, failed);
    }
    return hipGetLastError();
}

// --------------------------------------------------------------------------
// The results of a one-round grid solve in one block of device memory, so
// that they reach the host in one transfer: the inside count and the
// failed-walk count, then every per-point array (the inside points' arrays
// hold n_in of their n slots).
// --------------------------------------------------------------------------
struct PackLayout {
    size_t bval, mean, m2, cnt, ss, amb, hits, loc, total;
    __host__ __device__ explicit PackLayout(int64_t n) {
        const size_t m = (size_t)n;
        size_t o = 16;
        bval = o; o += 8 * m;
        mean = o; o += 8 * m;
        m2 = o; o += 8 * m;
        cnt = o; o += 8 * m;
        ss = o; o += 8 * m;
        amb = o; o += 4 * m;
        hits = o; o += 4 * m;
        loc = o; o += m;
        total = o;
    }
};

__global__ __launch_bounds__(kBlock) void k_pack_results(
    int64_t n, const int32_t *n_in_dev, const unsigned long long *failed, const uint8_t *loc,
    const int32_t *amb, const double *bval, const double *mean, const double *m2,
    const double *cnt, const double *ss, const int32_t *hits, unsigned char *pack) {
    const PackLayout L(n);
    const int64_t n_in = *n_in_dev;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        reinterpret_cast<int64_t *>(pack)[0] = n_in;
        reinterpret_cast<unsigned long long *>(pack)[1] = *failed;
    }
    if (i >= n) return;
    reinterpret_cast<double *>(pack + L.bval)[i] = bval[i];
    reinterpret_cast<int32_t *>(pack + L.amb)[i] = amb[i];
    pack[L.loc + (size_t)i] = loc[i];
    if (i < n_in) {
        reinterpret_cast<double *>(pack + L.mean)[i] = mean[i];
        reinterpret_cast<double *>(pack + L.m2)[i] = m2[i];
        reinterpret_cast<double *>(pack + L.cnt)[i] = cnt[i];
        reinterpret_cast<double *>(pack + L.ss)[i] = ss[i];
        reinterpret_cast<int32_t *>(pack + L.hits)[i] = hits[i];
    }
}

// --------------------------------------------------------------------------
// kernel 4 (adjoint): per point, sum the walks' parameter derivatives in walk
// order; scatter coef x (weight x blend) to the exit primitive's vertices and
// coef x weight to its boundary id.  coef = dL/du_p / N_p.
// --------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_adj_reduce(RoundView rv, AdjOut adj, MeshView m,
                                                       const double *coef, double *scal_sum,
                                                       double *grad_v, double *grad_bnd) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= rv.n) return;
    const int p = rv.pid[k];
    const int64_t base = rv.off ? rv.off[k] : (int64_t)k * rv.cw;
    const int cnt = rv.off ? (int)(rv.off[k + 1] - rv.off[k]) : rv.cw;
    const double c = coef[p];
    double s[6];
    for (int i = 0; i < 6; ++i) s[i] = scal_sum[(size_t)p * 6 + i];
    for (int j = 0; j < cnt; ++j) {
        const size_t o = (size_t)(base + j);
        for (int i = 0; i < 6; ++i) s[i] += adj.scal[o * 6 + i];
        const int prim = adj.prim[o];
        if (prim < 0 || c == 0.0) continue;
        const double *vw = adj.vw + o * 3;
        if (grad_v)
            for (int v = 0; v < m.dim; ++v)
                atomicAdd(&grad_v[m.prims[(size_t)prim * m.dim + v]], c * vw[v]);
        if (grad_bnd) {
            const int b = m.prim_bid[prim];
            if (b >= 0 && b < WOS_MAX_BND_VALUES) atomicAdd(&grad_bnd[b], c * (vw[0] + vw[1] + vw[2]));
        }
    }
    for (int i = 0; i < 6; ++i) scal_sum[(size_t)p * 6 + i] = s[i];
}

// boundary points: u = g(start's nearest point), linear in the boundary data
__global__ __launch_bounds__(kBlock) void k_adj_boundary(MeshView m, const uint8_t *loc,
                                                         const StartInfo *starts,
                                                         const double *dl_du, int64_t n,
                                                         double *grad_v, double *grad_bnd) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || loc[i] != 2) return;
    const StartInfo &S = starts[i];
    const double c = dl_du[i];
    if (S.prim < 0 || c == 0.0) return;
    if (grad_v) {
        const double t = m.dim == 3 ? S.t : 0.0;
        const double wv[3] = {1.0 - S.s - t, S.s, t};
        for (int v = 0; v < m.dim; ++v)
            atomicAdd(&grad_v[m.prims[(size_t)S.prim * m.dim + v]], c * wv[v]);
    }
    if (grad_bnd) {
        const int b = m.prim_bid[S.prim];
        if (b >= 0 && b < WOS_MAX_BND_VALUES) atomicAdd(&grad_bnd[b], c);
    }
}

// --------------------------------------------------------------------------
// host runtime
// --------------------------------------------------------------------------
thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

#define HIPX(expr)                                                                          \
    do {                                                                                    \
        hipError_t _e = (expr);                                                             \
        if (_e != hipSuccess)                                                               \
            return fail(WOS_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    ~DevBuf() { if (p) (void)hipFree(p); }
    hipError_t reserve(size_t n) {
        if (n <= bytes && p) return hipSuccess;
        if (p) { (void)hipFree(p); p = nullptr; bytes = 0; }
        hipError_t e = hipMalloc(&p, n ? n : 16);
        if (e == hipSuccess) bytes = n ? n : 16;
        return e;
    }
    template <typename T> T *as() const { return static_cast<T *>(p); }
};

// page-locked host memory (hipHostMalloc), grown on demand: device-to-host
// copies into it are single DMA transfers
struct PinnedBuf {
    void *p = nullptr;
    size_t bytes = 0;
    ~PinnedBuf() { if (p) (void)hipHostFree(p); }
    hipError_t reserve(size_t n) {
        if (n <= bytes && p) return hipSuccess;
        if (p) { (void)hipHostFree(p); p = nullptr; bytes = 0; }
        hipError_t e = hipHostMalloc(&p, n ? n : 16, hipHostMallocDefault);
        if (e == hipSuccess) bytes = n ? n : 16;
        return e;
    }
};

template <typename T>
hipError_t upload(DevBuf &b, const std::vector<T> &v, hipStream_t s) {
    hipError_t e = b.reserve(v.size() * sizeof(T));
    if (e != hipSuccess || v.empty()) return e;
    e = hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s);
}

}  // namespace

// Staging memory of wos_closest_points_noalloc: one fixed array in each
// GPU's copy of the code object, held until the process ends, so that the
// entry point allocates nothing.  In host code `scratch` is only a symbol:
// its device address, resolved per device by scratch_address(), is what
// kernels and copies use.  Every call on a GPU shares that GPU's array.
constexpr size_t kScratchBytes = (size_t)16 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_addr[kMaxDevices] = {};

// The device address of `scratch` on `device` (which must be current),
// resolved once per device.
hipError_t scratch_address(int device, void **out) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_addr[device] = p;
    }
    *out = g_scratch_addr[device];
    return hipSuccess;
}

}  // namespace

// what the autotuner's cached shapes are keyed by
struct TuneKey {
    int dim = 0, n_domains = 0;
    int64_t n_prims = 0, n_points = 0;
    int kind = 0, source_mode = 0, source_kind = 0, grid_cells = 0, mode = 0;
    bool operator==(const TuneKey &o) const {
        return dim == o.dim && n_domains == o.n_domains && n_prims == o.n_prims && mode == o.mode &&
               n_points == o.n_points && kind == o.kind && source_mode == o.source_mode &&
               source_kind == o.source_kind && grid_cells == o.grid_cells;
    }
};

struct wos_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int n_cu = 0;
    HostMesh mesh;
    DeviceLayouts L;
    bool has_mesh = false;
    DevBuf nd2, nd3, segd, trid, slot_prim, prim_bid, prims, verts, nf2, nf3, segf, trif;
    DevBuf nw2, nw3, doms_w;          // the 4-wide FP32 forest and its domain roots
    DevBuf cq_out, cq_prim;
    DevBuf deepq;                     // k_classify3's queue of points for k_classify3_deep
    DevBuf units;                     // FP32 query-unit counters
    DevBuf wstack;                    // FP32 walker traversal stacks (every lane of a full GPU)
    DevBuf failed;                    // walks with a NaN value (one counter)
    DevBuf pack;                      // one-round grid solve results, packed (PackLayout)
    DevBuf lane_st;                   // FP32 merge lane states of multi-round solves
    PinnedBuf hpack;                  // their host copy
    PinnedBuf hcls, hstat;            // multi-round solves: classification and statistics staging
    std::vector<std::pair<TuneKey, wos_tuning>> tuned;  // autotuned shapes
    // graph replay: the captured solve, what it was captured for, its tolerances
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    std::vector<uint8_t> graph_key;
    std::vector<double> graph_tol;
    int graph_blocks = 0;
    DevBuf g_bcount, g_boff, g_misc;
    bool count_units = false;
    DevBuf g_info, g_cells, g_list;   // distance grid
    DevBuf g_planes;                  // its planar cells' planes (3D)
    bool grid_planar = false;         // the grid has planar cells
    bool grid_valid = false;          // the grid below belongs to the current mesh
    int grid_built_cells = 0;         // the tuning's grid_cells it was built for
    int grid_built_cap = 0;           // its list cap
    float grid_built_relax = -1.0f;   // and the relax it was built for
    int64_t grid_list_len = 0;
    int64_t grid_bytes = 0;           // device bytes of cells, lists and grid records
    double grid_build_pending_s = 0.0;  // last build's seconds, until a summary reports them
    DevBuf doms;                      // DomainF per domain
    DevBuf starts_f, r_kblk;          // FP32 start copies; item block -> point (uneven rounds)
    DevBuf ref_state;                 // reference-stream state per inside point (8 words)
    DevBuf sidx;                      // per-point seed indices of an indexed point list
    DevBuf plist_in, loc, starts, amb, bvalue, inside, vals, steps, mean, m2, count, ssum, hits,
        counter, vv, pid_all, r_pid, r_wbeg, r_off, coef, scal_sum, grad_v, grad_b, dl_du,
        adj_scal, adj_vw, adj_prim;
    DevBuf field;                     // WOS_SRC_FIELD: the previous field
    DevBuf term_pw, term_slot;        // FP32 walks' end points and slots (g != 0)
    // an uneven round's point tables (off, pid, wbeg) on the host: two, so
    // that the next round's are written while the last upload may still be
    // queued; each is rewritten only after its upload's event
    PinnedBuf h_round[2];
    hipEvent_t h_round_ev[2]{};
    bool h_round_used[2]{};
    std::vector<hipEvent_t> round_ev;  // walker / merge times of queued rounds (3 a round)
    int last_rounds = 0;
    MeshView view{};
    hipEvent_t ev[6]{};
    // defaults: two wavefronts per block and 40 ready lanes per walk-step
    // phase measured 5.6 % faster than 256 and 48 on the icosphere benchmark
    // and no slower on the plate and star workloads
    wos_tuning tune{128, -1, 64, 0, 40, 0.2f, -1, 127, 8, 2, 0};
    size_t lds_per_cu = 65536;
    size_t lds_per_block = 65536;
    int last_threads = 0;
    int last_cached = 0;
    void *scratch_dev = nullptr;      // this device's `scratch` array (no-allocation entry point)

    wos_ctx() = default;
    wos_ctx(const wos_ctx &) = delete;
    wos_ctx &operator=(const wos_ctx &) = delete;
    // the HIP objects this context made (a failed wos_ctx_create included);
    // the device buffers are freed after this, as members
    ~wos_ctx() {
        (void)hipSetDevice(device);
        if (stream) (void)hipStreamSynchronize(stream);
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto &e : ev) if (e) (void)hipEventDestroy(e);
        for (auto &e : h_round_ev) if (e) (void)hipEventDestroy(e);
        for (auto &e : round_ev) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// Cells of a domain's distance grid.  grid_cells > 0: that many.  -1 (the
// default): sized from the domain's primitive count, 512 cells per segment
// (2D) or 4096 per triangle (3D) rounded up to a power of two, at least
// 2^16 and at most 2^22 (2D) or 2^24 (3D).  Measured: in 3D the walker runs
// faster with every doubling up to 2^24 cells (icosphere s4: 42.8 ms at
// 2^18, 31.0 at 2^22, 29.0 at 2^24; the 200k-triangle plate 420, 181, 147
// ms at 2^20, 2^22, 2^24) against a build of 17-150 ms once per mesh; in 2D
// per-domain sizing is fastest (16 star domains of 512-16384 segments:
// 187 ms against 197 ms at 2^22 cells each, with half the memory).
constexpr int64_t kGridMinCells = (int64_t)1 << 16;
int64_t grid_cells_for(int tune_cells, int dim, int64_t n_prims) {
    if (tune_cells >= 0) return tune_cells;
    const int64_t per_prim = dim == 3 ? 4096 : 512;
    const int64_t hi = dim == 3 ? (int64_t)1 << 24 : (int64_t)1 << 22;
    int64_t cells = kGridMinCells;
    while (cells < hi && cells < per_prim * n_prims) cells <<= 1;
    return cells;
}

// Build the distance grid of every domain for the FP32 walker (once per mesh
// and grid setting) and point the kernel view at it; grid_cells == 0 turns
// it off (see grid_cells_for for the sizes).  Cubic cells of size
// h = (bbox volume / cells)^(1/dim) over the domain's bounding box padded by
// half a cell.  The build's seconds are kept for the next solve summary.
hipError_t ensure_grid(wos_ctx *c) {
    if (c->tune.grid_cells == 0) {
        c->view.grids = nullptr; c->view.cells = nullptr; c->view.glist = nullptr;
        c->view.planes = nullptr;
        return hipSuccess;
    }
    if (!c->grid_valid || c->grid_built_cells != c->tune.grid_cells ||
        c->grid_built_cap != c->tune.grid_list_max || c->grid_built_relax != c->tune.relax) {
        hipError_t e0 = hipStreamSynchronize(c->stream);
        if (e0 != hipSuccess) return e0;
        const auto t0 = std::chrono::steady_clock::now();
        c->grid_valid = false;
        const int dim = c->mesh.dim;
        const int nd = (int)c->L.domains.size();
        std::vector<GridInfo> info(nd);
        int64_t total = 0;
        for (int d = 0; d < nd; ++d) {
            const DomainInfo &D = c->L.domains[d];
            const int64_t target = grid_cells_for(c->tune.grid_cells, dim, D.n_prims);
            double ext[3], emax = 0.0;
            for (int k = 0; k < dim; ++k) {
                ext[k] = D.bounds[2 * k + 1] - D.bounds[2 * k];
                emax = std::max(emax, ext[k]);
            }
            if (!(emax > 0.0)) emax = 1.0;
            double vol = 1.0;
            for (int k = 0; k < dim; ++k) vol *= std::max(ext[k], 1e-3 * emax);
            const double h = std::pow(vol / (double)target, 1.0 / dim);
            GridInfo &G = info[d];
            int32_t nk[3] = {1, 1, 1};
            for (int k = 0; k < dim; ++k)
                nk[k] = (int32_t)std::min<double>(std::floor(ext[k] / h) + 2.0, 1 << 15);
            G.nx = nk[0]; G.ny = nk[1]; G.nz = nk[2];
            G.h = (float)h;
            G.inv_h = 1.0f / G.h;
            G.ox = (float)(D.bounds[0] - 0.5 * h);
            G.oy = (float)(D.bounds[2] - 0.5 * h);
            G.oz = dim == 3 ? (float)(D.bounds[4] - 0.5 * h) : 0.0f;
            G.hd = (float)(0.5 * h * std::sqrt((double)dim));
            G.cell_off = (int32_t)total;
            total += (int64_t)G.nx * G.ny * G.nz;
            if (total > (int64_t)1 << 30) return hipErrorOutOfMemory;
        }
        DevBuf counts, rl2, offs, n_planar;
        hipError_t e;
        if ((e = n_planar.reserve(sizeof(unsigned int))) != hipSuccess) return e;
        if ((e = hipMemsetAsync(n_planar.p, 0, sizeof(unsigned int), c->stream)) != hipSuccess)
            return e;
        if ((e = c->g_cells.reserve((size_t)total * sizeof(GridCellF))) != hipSuccess) return e;
        if (dim == 3 && (e = c->g_planes.reserve((size_t)total * sizeof(float4))) != hipSuccess)
            return e;
        if ((e = counts.reserve((size_t)total * sizeof(int32_t))) != hipSuccess) return e;
        if ((e = rl2.reserve((size_t)total * sizeof(double))) != hipSuccess) return e;
        const size_t lds = (size_t)kBlock * (c->L.depth + 2) * sizeof(int);
        MeshView v = c->view;
        for (int phase = 0; phase < 2; ++phase) {
            if (phase == 1) {
                std::vector<int32_t> cnt((size_t)total);
                if ((e = hipMemcpyAsync(cnt.data(), counts.p, cnt.size() * sizeof(int32_t),
                                        hipMemcpyDeviceToHost, c->stream)) != hipSuccess) return e;
                if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return e;
                // every list starts at an odd entry index (lengths padded to
                // even), so that entry 1 and each later pair is one 16 B
                // load; one spare entry past the end for the last pair
                std::vector<int64_t> off((size_t)total);
                int64_t run = 1;
                for (size_t i = 0; i < cnt.size(); ++i) {
                    off[i] = run;
                    run += cnt[i] + (cnt[i] & 1);
                }
                if (run >= ((int64_t)1 << 31)) return hipErrorOutOfMemory;
                c->grid_list_len = run + 1;
                if ((e = upload(offs, off, c->stream)) != hipSuccess) return e;
                if ((e = c->g_list.reserve((size_t)(run + 1) * sizeof(GridEntry))) != hipSuccess)
                    return e;
                if ((e = hipMemsetAsync(c->g_list.p, 0, (size_t)(run + 1) * sizeof(GridEntry),
                                        c->stream)) != hipSuccess)
                    return e;
            }
            for (int d = 0; d < nd; ++d) {
                const GridInfo &G = info[d];
                const DomainInfo &D = c->L.domains[d];
                const int n = G.nx * G.ny * G.nz;
                const double hd = 0.5 * (double)G.h * std::sqrt((double)dim) * 1.0001 +
                                  64.0 * std::ldexp(1.0, -24) * D.coord_scale;
                const double pad = 2.0 * hd;
                // with relax r, a cell with dc >= hd (2 - r) / r answers every
                // query from its bound; lists are kept below that (padded)
                const double r = (double)c->tune.relax;
                const double band = r > 0.0 ? 1.05 * hd * (2.0 - r) / r + pad : INFINITY;
                const int cap = c->tune.grid_list_max;
                const int blocks = (n + kBlock - 1) / kBlock;
                // coplanar within a quarter of an FP32 ulp of the domain's
                // coordinates: the plane distance's own FP32 error (a few
                // ulp) stays inside the walker's radius margin (32 ulp)
                const double plane_tol = std::ldexp(1.0, -26) * D.coord_scale;
                if (dim == 2)
                    hipLaunchKernelGGL(k_grid_build<2>, dim3(blocks), dim3(kBlock), lds, c->stream, v,
                                       G, D.node_off, pad, cap, band, phase, c->g_cells.as<GridCellF>(),
                                       counts.as<int32_t>(), rl2.as<double>(), offs.as<int64_t>(),
                                       c->g_list.as<GridEntry>(), nullptr, 0.0, nullptr);
                else
                    hipLaunchKernelGGL(k_grid_build<3>, dim3(blocks), dim3(kBlock), lds, c->stream, v,
                                       G, D.node_off, pad, cap, band, phase, c->g_cells.as<GridCellF>(),
                                       counts.as<int32_t>(), rl2.as<double>(), offs.as<int64_t>(),
                                       c->g_list.as<GridEntry>(), c->g_planes.as<float4>(),
                                       plane_tol, n_planar.as<unsigned int>());
                if ((e = hipGetLastError()) != hipSuccess) return e;
            }
        }
        unsigned int planar_cells = 0;
        if ((e = hipMemcpyAsync(&planar_cells, n_planar.p, sizeof(planar_cells),
                                hipMemcpyDeviceToHost, c->stream)) != hipSuccess)
            return e;
        if ((e = upload(c->g_info, info, c->stream)) != hipSuccess) return e;
        if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return e;
        c->grid_planar = planar_cells > 0;
        c->grid_built_cells = c->tune.grid_cells;
        c->grid_built_cap = c->tune.grid_list_max;
        c->grid_built_relax = c->tune.relax;
        c->grid_bytes = total * (int64_t)sizeof(GridCellF) +
                        (dim == 3 ? total * (int64_t)sizeof(float4) : 0) +
                        c->grid_list_len * (int64_t)sizeof(GridEntry) +
                        nd * (int64_t)sizeof(GridInfo);
        c->grid_build_pending_s =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        c->grid_valid = true;
    }
    c->view.grids = c->g_info.as<GridInfo>();
    c->view.cells = c->g_cells.as<GridCellF>();
    c->view.glist = c->g_list.as<GridEntry>();
    c->view.planes = c->mesh.dim == 3 ? c->g_planes.as<float4>() : nullptr;
    return hipSuccess;
}

// Global traversal stacks of the FP32 walker: depth + 1 entries for every
// lane the GPU can hold (2048 per CU), sized once per mesh, before any
// graph capture.
// Stack entries a traversal needs: a binary node pushes at most one entry
// per level (depth + 1), a 4-wide node at most three (3 depth_w + 2).
int stack_entries(const wos_ctx *c, int width) {
    return width == 4 ? 3 * c->L.depth_w + 2 : c->L.depth + 1;
}

hipError_t ensure_walk_stack(wos_ctx *c) {
    const int per_lane = std::max(stack_entries(c, 2), stack_entries(c, 4));
    const size_t need = (size_t)c->n_cu * 2048 * per_lane * sizeof(StackEntry);
    return c->wstack.reserve(need);
}

template <typename K>
hipError_t launch_classify(wos_ctx *c, K kernel, const EqView &ev, const GridSpec &g,
                           const DomainView &dv, const double *plist, const int64_t *sidx,
                           int64_t n, uint64_t seed, int mra, int per_block, int threads) {
    const int blocks = (int)((n + per_block - 1) / per_block);
    const size_t lds = (size_t)threads * (c->L.depth + 2) * sizeof(int);
    hipLaunchKernelGGL(kernel, dim3(blocks), dim3(threads), lds, c->stream, c->view, ev, g, dv,
                       plist, sidx, n, seed, mra, c->loc.as<uint8_t>(), c->starts.as<StartInfo>(),
                       c->amb.as<int32_t>(), c->bvalue.as<double>());
    return hipGetLastError();
}

// end-point buffers of the deferred terminal projection (FP32 forward walks
// with g != 0), one entry per walk of a round
hipError_t reserve_terminal(wos_ctx *c, const EqView &ev, int precision, bool adjoint,
                                   int64_t total) {
    if (precision != WOS_FP32_PERSISTENT || adjoint || ev.boundary_kind == WOS_BND_ZERO)
        return hipSuccess;
    hipError_t e = c->term_pw.reserve((size_t)total * sizeof(float4));
    if (e != hipSuccess) return e;
    return c->term_slot.reserve((size_t)total * sizeof(int32_t));
}

template <int DIM, bool SRC, bool SCR, int SX>
int launch_ref_sx(wos_ctx *c, const EqView &ev, const RoundView &rv, double eps, int max_steps,
                  size_t lds) {
    // persistent: at most the blocks the GPU keeps resident
    const auto kfn = k_walk_ref<DIM, SRC, SCR, SX>;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, kBlock, lds) != hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    const int64_t need = ((int64_t)rv.n + kBlock - 1) / kBlock;
    const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>(need, (int64_t)per_cu * c->n_cu));
    unsigned long long *next = c->counter.as<unsigned long long>();
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, c->stream, next, 1);
    hipLaunchKernelGGL(kfn, dim3(blocks), dim3(kBlock), lds, c->stream, c->view, ev, rv, eps,
                       max_steps, c->vals.as<double>(), c->steps.as<int32_t>(),
                       c->ref_state.as<uint64_t>(), next, kFp64MinReady);
    return blocks;
}

template <int DIM, bool SRC, bool SCR, bool ADJ, int SX>
int launch_fp64_sx(wos_ctx *c, const EqView &ev, const RoundView &rv, double eps, int max_steps,
                   const AdjOut &adj, size_t lds) {
    // persistent: at most the blocks the GPU keeps resident
    const auto kfn = k_walk_fp64<DIM, SRC, SCR, ADJ, SX>;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, kBlock, lds) != hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    const int64_t need = (rv.total + kBlock - 1) / kBlock;
    const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>(need, (int64_t)per_cu * c->n_cu));
    unsigned long long *next = c->counter.as<unsigned long long>();
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, c->stream, next, 1);
    hipLaunchKernelGGL(kfn, dim3(blocks), dim3(kBlock), lds, c->stream, c->view, ev, rv, eps,
                       max_steps, c->vals.as<double>(), c->steps.as<int32_t>(), adj, next,
                       kFp64MinReady);
    return blocks;
}

// The FP32 persistent walker on the binary (W = 2) or 4-wide (W = 4) FP32
// forest: launch shape from the occupancy, LDS layout, arguments.
template <int DIM, int SMODE, int kM3, int kM4, int kM5, bool SCR, bool ADJ, int W, bool PL>
hipError_t launch_fp32_w(wos_ctx *c, const EqView &ev, int mode, const RoundView &rv, double eps,
                         int max_steps, const AdjOut &adj, int *blocks_out) {
    const void *kfn = mode == 3   ? (const void *)k_walk_fp32<DIM, kM3, SCR, ADJ, W, PL>
                      : mode == 4 ? (const void *)k_walk_fp32<DIM, kM4, SCR, ADJ, W, PL>
                      : mode == 5 ? (const void *)k_walk_fp32<DIM, kM5, SCR, ADJ, W, PL>
                      : mode == 2 ? (const void *)k_walk_fp32<DIM, 2, SCR, ADJ, W, PL>
                                  : (const void *)k_walk_fp32<DIM, SMODE, SCR, ADJ, W, PL>;
    const int threads = c->tune.block_threads;
    // LDS per block beyond the node cache: one triangle record per thread
    // (3D) and the first klds entries of every thread's traversal stack
    const int need = stack_entries(c, W);
    const int klds = std::min(c->tune.stack_lds, need);
    const size_t tri_bytes = DIM == 3 ? (size_t)threads * 3 * 16 + 16 : 0;
    const size_t lstack_bytes = (size_t)threads * klds * sizeof(StackEntry);
    const size_t stack_bytes = tri_bytes + lstack_bytes;
    const int node_bytes = W == 4 ? (DIM == 3 ? (int)sizeof(NodeF3W) : (int)sizeof(NodeF2W))
                                  : (DIM == 3 ? (int)sizeof(NodeF3) : (int)sizeof(NodeF2));
    const int n_internal = W == 4 ? (DIM == 3 ? (int)c->L.nw3.size() : (int)c->L.nw2.size())
                                  : (DIM == 3 ? (int)c->L.nf3.size() : (int)c->L.nf2.size());
    // occupancy allowed by registers alone, then the LDS each block may use
    // for the node cache without lowering it
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, threads, stack_bytes);
    if (e != hipSuccess) return e;
    per_cu = std::max(1, per_cu);
    if (c->tune.blocks_per_cu > 0) per_cu = std::min(per_cu, c->tune.blocks_per_cu);
    int n_cached = c->tune.lds_nodes;
    if (n_cached < 0) {
        const size_t per_block = std::min<size_t>(c->lds_per_cu / per_cu, c->lds_per_block);
        n_cached = per_block > stack_bytes ? (int)((per_block - stack_bytes) / node_bytes) : 0;
    }
    n_cached = std::max(0, std::min(n_cached, n_internal));
    const int cache_bytes = (n_cached * node_bytes + 15) / 16 * 16;
    const size_t lds = stack_bytes + cache_bytes + 16;
    if (lds > c->lds_per_block) return hipErrorInvalidValue;
    int per_cu2 = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu2, kfn, threads, lds);
    if (e != hipSuccess) return e;
    per_cu2 = std::max(1, std::min(per_cu2, per_cu));
    const int blocks = c->n_cu * per_cu2;
    if ((size_t)blocks * threads * need * sizeof(StackEntry) > c->wstack.bytes)
        return hipErrorInvalidValue;  // ensure_walk_stack sizes it for any launch
    PersistArgs a;
    a.gstack = c->wstack.as<StackEntry>();
    a.gstride = blocks * threads;
    a.klds = klds;
    a.n_cached = n_cached;
    a.cache_bytes = cache_bytes;
    // after the node cache: the triangle rows (16 B aligned), then the
    // LDS part of the stacks
    a.tri_off = DIM == 3 ? (int)(cache_bytes + 15) / 16 * 16 : -1;
    a.stack_off = (int)((cache_bytes + tri_bytes + 15) / 16 * 16);
    if ((size_t)a.stack_off + lstack_bytes > lds) return hipErrorInvalidValue;
    a.rv = rv;
    a.max_steps = max_steps;
    a.eps = eps;
    a.eps_f = (float)eps;
    a.shrink_rel = 1.0f - 1.0f / 131072.0f;
    a.cf = (1.0f - c->tune.relax) * (1.0f - c->tune.relax);
    a.min_ready = c->tune.min_ready;
    // refilling once 8 lanes are free rather than at every free lane runs
    // the refill code (item lookup, start record, domain record) about a
    // third as often; measured 1.4-2 % faster walks than 1 or 16 on the
    // icosphere, plate and star workloads, with identical results
    a.refill_min = 8;
    a.counter = c->counter.as<unsigned long long>();
    a.chunk = c->tune.chunk;
    a.vals = c->vals.as<double>();
    a.steps_out = c->steps.as<int32_t>();
    a.units = c->count_units ? c->units.as<unsigned long long>() : nullptr;
    const bool defer = !ADJ && ev.boundary_kind != WOS_BND_ZERO;
    a.term_pw = defer ? c->term_pw.as<float4>() : nullptr;
    a.term_slot = defer ? c->term_slot.as<int32_t>() : nullptr;
    a.parts = c->tune.xcd_parts;
    // the 4-wide forest has its own domain roots
    MeshView view = c->view;
    if (W == 4) view.doms = c->doms_w.as<DomainF>();
    // the cold record sits after the part counters in the same buffer
    WalkCold cold{a.rv, a.counter, a.vals, a.steps_out, a.term_pw, a.term_slot, a.parts, a.chunk};
    WalkCold *cold_dst = reinterpret_cast<WalkCold *>(a.counter + kCounterWords);
    a.cold = cold_dst;
    hipLaunchKernelGGL(k_init_parts, dim3(1), dim3(64), 0, c->stream, a.counter, a.parts,
                       (int64_t)rv.total, rv.total_dev, cold, cold_dst);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (mode == 3)
        hipLaunchKernelGGL((k_walk_fp32<DIM, kM3, SCR, ADJ, W, PL>), dim3(blocks), dim3(threads), lds,
                           c->stream, view, ev, a, adj);
    else if (mode == 4)
        hipLaunchKernelGGL((k_walk_fp32<DIM, kM4, SCR, ADJ, W, PL>), dim3(blocks), dim3(threads), lds,
                           c->stream, view, ev, a, adj);
    else if (mode == 5)
        hipLaunchKernelGGL((k_walk_fp32<DIM, kM5, SCR, ADJ, W, PL>), dim3(blocks), dim3(threads), lds,
                           c->stream, view, ev, a, adj);
    else if (mode == 2)
        hipLaunchKernelGGL((k_walk_fp32<DIM, 2, SCR, ADJ, W, PL>), dim3(blocks), dim3(threads), lds,
                           c->stream, view, ev, a, adj);
    else
        hipLaunchKernelGGL((k_walk_fp32<DIM, SMODE, SCR, ADJ, W, PL>), dim3(blocks), dim3(threads), lds,
                           c->stream, view, ev, a, adj);
    *blocks_out = blocks;
    c->last_threads = threads;
    c->last_cached = n_cached;
    return hipGetLastError();
}

template <int DIM, bool SRC, bool SCR, bool ADJ>
hipError_t launch_walk(wos_ctx *c, const EqView &ev, int precision, const RoundView &rv,
                       double eps, int max_steps, const AdjOut &adj, int *blocks_out) {
    // source extensions, 2D forward solves only (the descriptor checks and
    // the adjoint's refusals keep every other case off their kernels):
    // 1 GREEN_MIS, 2 WOS_SRC_FIELD
    constexpr bool kExt = SRC && DIM == 2 && !ADJ;
    constexpr int kMis = kExt ? 1 : 0, kFld = kExt ? 2 : 0;
    const int sx = !kExt ? 0
                   : ev.source_mode == WOS_SOURCE_GREEN_MIS ? 1
                   : ev.source_kind == WOS_SRC_FIELD      ? 2
                                                          : 0;
    if (precision == WOS_FP64_REFERENCE) {
        if constexpr (ADJ) {
            return hipErrorInvalidValue;
        } else {
            const size_t lds = (size_t)kBlock * (c->L.depth + 2) * sizeof(int);
            *blocks_out = sx == 1   ? launch_ref_sx<DIM, SRC, SCR, kMis>(c, ev, rv, eps, max_steps, lds)
                          : sx == 2 ? launch_ref_sx<DIM, SRC, SCR, kFld>(c, ev, rv, eps, max_steps, lds)
                                    : launch_ref_sx<DIM, SRC, SCR, 0>(c, ev, rv, eps, max_steps, lds);
            return hipGetLastError();
        }
    }
    if (precision == WOS_FP64_PLAIN) {
        const size_t lds = (size_t)kBlock * (c->L.depth + 2) * sizeof(int);
        *blocks_out = sx == 1   ? launch_fp64_sx<DIM, SRC, SCR, ADJ, kMis>(c, ev, rv, eps, max_steps, adj, lds)
                      : sx == 2 ? launch_fp64_sx<DIM, SRC, SCR, ADJ, kFld>(c, ev, rv, eps, max_steps, adj, lds)
                                : launch_fp64_sx<DIM, SRC, SCR, ADJ, 0>(c, ev, rv, eps, max_steps, adj, lds);
        return hipGetLastError();
    }
    hipError_t ge = ensure_grid(c);
    if (ge != hipSuccess) return ge;
    // FP32 source modes: 0 none, 1 uniform, 2 Green, 3 GREEN_MIS, 4 Green
    // with the field source, 5 uniform with the field source
    constexpr int SMODE = SRC ? 1 : 0;
    constexpr int kM3 = kExt ? 3 : SMODE, kM4 = kExt ? 4 : SMODE, kM5 = kExt ? 5 : SMODE;
    const bool green = ev.source_mode == WOS_SOURCE_GREEN && SRC;
    const int mode = sx == 1 ? 3 : sx == 2 ? (green ? 4 : 5) : green ? 2 : SMODE;
    // the planar-cell query only where the grid has planar cells (3D): the
    // other kernels do without its test
    if constexpr (DIM == 3) {
        if (c->grid_planar && c->view.planes) {
            if (c->tune.bvh_width == 4)
                return launch_fp32_w<DIM, SMODE, kM3, kM4, kM5, SCR, ADJ, 4, true>(
                    c, ev, mode, rv, eps, max_steps, adj, blocks_out);
            return launch_fp32_w<DIM, SMODE, kM3, kM4, kM5, SCR, ADJ, 2, true>(
                c, ev, mode, rv, eps, max_steps, adj, blocks_out);
        }
    }
    if (c->tune.bvh_width == 4)
        return launch_fp32_w<DIM, SMODE, kM3, kM4, kM5, SCR, ADJ, 4, false>(c, ev, mode, rv, eps,
                                                                             max_steps, adj, blocks_out);
    return launch_fp32_w<DIM, SMODE, kM3, kM4, kM5, SCR, ADJ, 2, false>(c, ev, mode, rv, eps,
                                                                         max_steps, adj, blocks_out);
}

template <int DIM, bool ADJ>
hipError_t dispatch_walk_t(wos_ctx *c, const EqView &ev, bool src, bool scr, int precision,
                           const RoundView &rv, double eps, int max_steps, const AdjOut &adj,
                           int *blocks) {
    if (src && scr) return launch_walk<DIM, true, true, ADJ>(c, ev, precision, rv, eps, max_steps, adj, blocks);
    if (src) return launch_walk<DIM, true, false, ADJ>(c, ev, precision, rv, eps, max_steps, adj, blocks);
    if (scr) return launch_walk<DIM, false, true, ADJ>(c, ev, precision, rv, eps, max_steps, adj, blocks);
    return launch_walk<DIM, false, false, ADJ>(c, ev, precision, rv, eps, max_steps, adj, blocks);
}

hipError_t dispatch_walk(wos_ctx *c, int dim, bool adjoint, const EqView &ev, bool src, bool scr,
                         int precision, const RoundView &rv, double eps, int max_steps,
                         const AdjOut &adj, int *blocks) {
    if (dim == 2)
        return adjoint ? dispatch_walk_t<2, true>(c, ev, src, scr, precision, rv, eps, max_steps, adj, blocks)
                       : dispatch_walk_t<2, false>(c, ev, src, scr, precision, rv, eps, max_steps, adj, blocks);
    return adjoint ? dispatch_walk_t<3, true>(c, ev, src, scr, precision, rv, eps, max_steps, adj, blocks)
                   : dispatch_walk_t<3, false>(c, ev, src, scr, precision, rv, eps, max_steps, adj, blocks);
}

// Width and normalisation of the GREEN_MIS proposal: the source
// exp(-|x - c|^2 / R^2) is a normal of per-axis width sigma = R / sqrt 2, and
// its integral over the box (limits +-inf without one) is the product over
// both axes of sigma sqrt(pi / 2) (erf((hi - c) / (sqrt 2 sigma)) -
// erf((lo - c) / (sqrt 2 sigma))), gaussian_heat_source.hpp:116-126.
double proposal_sigma(const wos_eq_desc *eq) { return eq->src_radius / std::sqrt(2.0); }
bool proposal_boxed(const wos_eq_desc *eq) {
    const double *b = eq->proposal_box;
    return !(b[0] == 0.0 && b[1] == 0.0 && b[2] == 0.0 && b[3] == 0.0);
}
double proposal_norm(const wos_eq_desc *eq) {
    constexpr double sqrt_two = 1.41421356237309504880168872420969808;
    constexpr double sqrt_pi_over_two = 1.25331413731550025120788264240552263;
    const double sigma = proposal_sigma(eq);
    const bool boxed = proposal_boxed(eq);
    double norm = 1.0;
    for (int k = 0; k < 2; ++k) {
        const double lo = boxed ? eq->proposal_box[2 * k] : -INFINITY;
        const double hi = boxed ? eq->proposal_box[2 * k + 1] : INFINITY;
        const double c = eq->src_center[k];
        norm *= sigma * sqrt_pi_over_two *
                (std::erf((hi - c) / (sqrt_two * sigma)) - std::erf((lo - c) / (sqrt_two * sigma)));
    }
    return norm;
}

int validate_eq(const wos_ctx *c, const wos_eq_desc *eq) {
    if (!eq) return fail(WOS_ERR_INPUT, "equation descriptor is NULL");
    if (eq->kind < 0 || eq->kind > 3) return fail(WOS_ERR_INPUT, "unknown equation kind");
    if (eq->source_mode < 0 || eq->source_mode > 2) return fail(WOS_ERR_INPUT, "unknown source mode");
    if (eq->source_kind < 0 || eq->source_kind > 6) return fail(WOS_ERR_INPUT, "unknown source kind");
    if (eq->boundary_kind < 0 || eq->boundary_kind > 6) return fail(WOS_ERR_INPUT, "unknown boundary kind");
    if (eq->kind == WOS_EQ_LAPLACE && eq->source_kind != WOS_SRC_NONE)
        return fail(WOS_ERR_INPUT, "laplace takes no source");
    if (eq->kind == WOS_EQ_SCREENED && !(eq->alpha > 0.0 && std::isfinite(eq->alpha)))
        return fail(WOS_ERR_INPUT, "screened_poisson needs a positive finite alpha");
    if (eq->kind == WOS_EQ_HELMHOLTZ && eq->source_kind != WOS_SRC_NONE)
        return fail(WOS_ERR_INPUT, "helmholtz takes no source");
    if (eq->kind == WOS_EQ_HELMHOLTZ && !(eq->alpha > 0.0 && std::isfinite(eq->alpha)))
        return fail(WOS_ERR_INPUT, "helmholtz needs a positive finite wavenumber (alpha)");
    if (eq->source_kind == WOS_SRC_GAUSSIAN && !(eq->src_radius > 0.0))
        return fail(WOS_ERR_INPUT, "gaussian source needs a positive width");
    if (eq->source_mode == WOS_SOURCE_GREEN_MIS && eq->source_kind != WOS_SRC_NONE) {
        // source_sampling.hpp:43-47 and wos.hpp:89-92; the proposal's own
        // checks follow gaussian_heat_source.hpp:29-35
        if (c->mesh.dim != 2 || eq->source_kind != WOS_SRC_GAUSSIAN ||
            (eq->kind != WOS_EQ_POISSON && eq->kind != WOS_EQ_SCREENED))
            return fail(WOS_ERR_INPUT, "GreenSourceMIS requires a source proposal sampler and PDF "
                                       "(a 2D Gaussian source of poisson or screened_poisson)");
        const double beta = eq->mis_green_probability;
        if (!(beta > 0.0 && beta < 1.0) || !std::isfinite(beta))
            return fail(WOS_ERR_INPUT, "Green-source MIS probability must be finite and in (0, 1)");
        const double *b = eq->proposal_box;
        const bool none = b[0] == 0.0 && b[1] == 0.0 && b[2] == 0.0 && b[3] == 0.0;
        if (!none && !(b[1] > b[0] && b[3] > b[2] && eq->src_center[0] >= b[0] &&
                       eq->src_center[0] <= b[1] && eq->src_center[1] >= b[2] &&
                       eq->src_center[1] <= b[3]))
            return fail(WOS_ERR_INPUT, "invalid Gaussian source proposal box");
        if (!(proposal_norm(eq) > 0.0) || !std::isfinite(proposal_norm(eq)))
            return fail(WOS_ERR_INPUT, "Gaussian source proposal has no mass in its box");
    }
    if (eq->source_kind == WOS_SRC_FIELD) {
        const wos_field_source *f = eq->field_source;
        if (c->mesh.dim != 2) return fail(WOS_ERR_INPUT, "the field source is 2D");
        if (!f) return fail(WOS_ERR_INPUT, "WOS_SRC_FIELD needs field_source");
        if (eq->source_mode == WOS_SOURCE_GREEN_MIS)
            return fail(WOS_ERR_INPUT, "GreenSourceMIS requires a source proposal sampler and PDF "
                                       "(a 2D Gaussian source of poisson or screened_poisson)");
        if (f->values) {
            // field_sampler.hpp:38-49
            if (f->nx < 2 || f->ny < 2)
                return fail(WOS_ERR_INPUT, "FieldSampler2D requires a complete grid of at least 2 x 2");
            if (!(f->bounds[1] > f->bounds[0]) || !(f->bounds[3] > f->bounds[2]))
                return fail(WOS_ERR_INPUT, "FieldSampler2D requires non-degenerate grid bounds");
            if (f->interpolation != WOS_INTERP_BILINEAR && f->interpolation != WOS_INTERP_CUBIC)
                return fail(WOS_ERR_INPUT, "unknown field interpolation");
            if ((int64_t)f->nx * f->ny > ((int64_t)1 << 28))
                return fail(WOS_ERR_INPUT, "field grid too large");
        }
        if (!std::isfinite(f->history_scale) || !std::isfinite(f->heat_scale))
            return fail(WOS_ERR_INPUT, "field source scales must be finite");
        if (f->n_heat_terms < 0 || f->n_heat_terms > WOS_MAX_HEAT_TERMS)
            return fail(WOS_ERR_INPUT, "n_heat_terms out of range");
        if (f->n_heat_terms > 0 && !(f->heat_divisor != 0.0 && std::isfinite(f->heat_divisor)))
            return fail(WOS_ERR_INPUT, "heat_divisor must be finite and nonzero");
        for (int k = 0; k < f->n_heat_terms; ++k) {
            const wos_heat_term &h = f->heat[k];
            if (!(h.sx > 0.0) || !(h.sy > 0.0) || !(h.b != 0.0) || !std::isfinite(h.a) ||
                !std::isfinite(h.b) || !std::isfinite(h.cx) || !std::isfinite(h.cy))
                return fail(WOS_ERR_INPUT, "invalid heat term");
        }
    }
    if (eq->n_bnd_values < 0 || eq->n_bnd_values > WOS_MAX_BND_VALUES)
        return fail(WOS_ERR_INPUT, "n_bnd_values out of range");
    if (eq->boundary_kind == WOS_BND_VERTEX &&
        (!eq->vertex_values || eq->n_vertex_values != c->mesh.n_verts()))
        return fail(WOS_ERR_INPUT, "per-vertex boundary values must cover every mesh vertex");
    return WOS_OK;
}

double elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, a, b);
    return ms * 1e-3;
}

// adjoint request: dL/du per point in, gradients out (any output may be NULL)
struct AdjRequest {
    const double *dl_du;
    double *grad_vertex;
    double *grad_params;
};

// shared body of wos_solve_grid / wos_solve_points / the adjoint entry points
// one slice of a solve: a grid over a domain's bbox or a point list in it
struct Part {
    int domain;
    const GridSpec *grid;
    const double *points;
    int64_t count;
    const int64_t *seed_index = nullptr;  // point list: per-point seed indices (else the list index)
};

// Validate a solve, build its equation view and stage its inputs.
int prepare_solve(wos_ctx *c, const std::vector<Part> &parts, const wos_eq_desc *eq,
                  const wos_run_params *prm, const AdjRequest *adjr, EqView &ev,
                  int64_t &n_out) {
    int64_t n = 0;
    for (const Part &p : parts) n += p.count;
    if (!c) return fail(WOS_ERR_INPUT, "context is NULL");
    if (!c->has_mesh) return fail(WOS_ERR_STATE, "no mesh set on this context");
    if (!prm) return fail(WOS_ERR_INPUT, "params are required");
    if (adjr && !adjr->dl_du) return fail(WOS_ERR_INPUT, "dl_du is required");
    if (adjr && prm->target_se > 0.0)
        return fail(WOS_ERR_INPUT, "the adjoint replays a fixed number of walks (target_se = 0)");
    if (adjr && prm->precision == WOS_FP64_REFERENCE)
        return fail(WOS_ERR_INPUT, "the adjoint replays counter streams (precision FP64 or FP32)");
    int rc = validate_eq(c, eq);
    if (rc) return rc;
    if (adjr && eq->source_mode == WOS_SOURCE_GREEN_MIS && eq->source_kind != WOS_SRC_NONE)
        return fail(WOS_ERR_INPUT, "the adjoint supports the uniform and Green source modes");
    if (adjr && eq->source_kind == WOS_SRC_FIELD)
        return fail(WOS_ERR_INPUT, "the adjoint does not differentiate field sources");
    if (prm->n_walks < 1) return fail(WOS_ERR_INPUT, "n_walks must be >= 1");
    if (prm->max_steps < 1) return fail(WOS_ERR_INPUT, "max_steps must be >= 1");
    if (c->mesh.dim == 3 && prm->max_ray_attempts < 3)
        return fail(WOS_ERR_INPUT, "max_ray_attempts must be >= 3");
    if (!(prm->epsilon > 0.0) || !std::isfinite(prm->epsilon))
        return fail(WOS_ERR_INPUT, "epsilon must be positive and finite");
    if (prm->precision != WOS_FP64_PLAIN && prm->precision != WOS_FP32_PERSISTENT &&
        prm->precision != WOS_FP64_REFERENCE)
        return fail(WOS_ERR_INPUT, "unknown precision");
    if (n <= 0 || n > ((int64_t)1 << 31) - 1) return fail(WOS_ERR_INPUT, "point count out of range");
    if (!(prm->target_se >= 0.0) || !std::isfinite(prm->target_se))
        return fail(WOS_ERR_INPUT, "target_se must be finite and >= 0");
    if (prm->target_se > 0.0 && prm->max_walks < prm->n_walks)
        return fail(WOS_ERR_INPUT, "adaptive mode needs max_walks >= n_walks");
    const int dim = c->mesh.dim;
    HIPX(hipSetDevice(c->device));

    ev = EqView{};
    ev.kind = eq->kind; ev.source_mode = eq->source_mode; ev.source_kind = eq->source_kind;
    ev.boundary_kind = eq->boundary_kind; ev.alpha = eq->alpha; ev.src_value = eq->src_value;
    ev.cx = eq->src_center[0]; ev.cy = eq->src_center[1]; ev.cz = eq->src_center[2];
    ev.src_radius = eq->src_radius;
    for (int i = 0; i < WOS_MAX_BND_VALUES; ++i) ev.bnd[i] = eq->bnd_values[i];
    ev.n_bnd = eq->n_bnd_values;
    if (eq->source_mode == WOS_SOURCE_GREEN_MIS && eq->source_kind != WOS_SRC_NONE) {
        ev.mis_beta = eq->mis_green_probability;
        ev.mis_sigma = proposal_sigma(eq);
        ev.mis_norm = proposal_norm(eq);
        ev.mis_box = proposal_boxed(eq) ? 1 : 0;
        for (int k = 0; k < 4; ++k) ev.box[k] = eq->proposal_box[k];
    }
    if (eq->source_kind == WOS_SRC_FIELD) {
        const wos_field_source *f = eq->field_source;
        ev.field = nullptr;
        if (f->values) {
            std::vector<double> fv(f->values, f->values + (size_t)f->nx * f->ny);
            HIPX(upload(c->field, fv, c->stream));
            ev.field = c->field.as<double>();
        }
        ev.fnx = f->nx; ev.fny = f->ny;
        ev.finterp = f->interpolation; ev.fclamp = f->clamp_nonnegative ? 1 : 0;
        for (int k = 0; k < 4; ++k) ev.fb[k] = f->bounds[k];
        ev.hist_scale = f->history_scale;
        ev.heat_scale = f->heat_scale;
        ev.heat_div = f->heat_divisor;
        ev.n_heat = f->n_heat_terms;
        for (int k = 0; k < f->n_heat_terms; ++k) ev.heat[k] = f->heat[k];
    }
    ev.vertex_values = nullptr;
    if (eq->boundary_kind == WOS_BND_VERTEX) {
        std::vector<double> vv(eq->vertex_values, eq->vertex_values + eq->n_vertex_values);
        HIPX(upload(c->vv, vv, c->stream));
        ev.vertex_values = c->vv.as<double>();
    }

    HIPX(c->failed.reserve(sizeof(unsigned long long)));
    HIPX(c->loc.reserve(n));
    HIPX(c->starts.reserve(n * sizeof(StartInfo)));
    HIPX(c->amb.reserve(n * sizeof(int32_t)));
    HIPX(c->bvalue.reserve(n * sizeof(double)));
    {
        std::vector<double> pts;
        for (const Part &p : parts)
            if (p.points) pts.insert(pts.end(), p.points, p.points + p.count * dim);
        if (!pts.empty()) HIPX(upload(c->plist_in, pts, c->stream));
        std::vector<int64_t> sidx;
        for (const Part &p : parts)
            if (p.points && p.seed_index) sidx.insert(sidx.end(), p.seed_index, p.seed_index + p.count);
        if (!sidx.empty()) HIPX(upload(c->sidx, sidx, c->stream));
    }

    n_out = n;
    return WOS_OK;
}

// Classify every part's points into its slice of loc/starts/amb/bvalue
// (stream-ordered, no synchronisation); part_tol receives the tolerances.
// conservative FP32 radius margin of a domain: 32 ulp of its largest coordinate
float fp32_margin(const DomainInfo &D) {
    return (float)(32.0 * std::ldexp(1.0, -24) * D.coord_scale);
}

// k_classify3<S> (64 points a block), then k_classify3_deep over the points
// it queued
template <int S>
hipError_t launch_classify3(wos_ctx *c, const EqView &ev, const GridSpec &g, const DomainView &dv,
                            const double *plist, const int64_t *sidx, int64_t n, uint64_t seed,
                            int mra) {
    hipError_t e = c->deepq.reserve(sizeof(unsigned long long) +
                                    (size_t)n * (sizeof(int64_t) + sizeof(int32_t)));
    if (e != hipSuccess) return e;
    unsigned char *base = c->deepq.as<unsigned char>();
    DeepQueue dq{reinterpret_cast<unsigned long long *>(base),
                 reinterpret_cast<int64_t *>(base + sizeof(unsigned long long)),
                 reinterpret_cast<int32_t *>(base + sizeof(unsigned long long) +
                                             (size_t)n * sizeof(int64_t))};
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, c->stream, dq.count, 1);
    const int threads = (S + 3) * 64;
    const int blocks = (int)((n + 63) / 64);
    size_t lds = (size_t)threads * (c->L.depth + 2) * sizeof(int);
    hipLaunchKernelGGL(k_classify3<S>, dim3(blocks), dim3(threads), lds, c->stream, c->view, ev, g,
                       dv, plist, sidx, n, seed, mra, c->loc.as<uint8_t>(),
                       c->starts.as<StartInfo>(), c->amb.as<int32_t>(), c->bvalue.as<double>(), dq);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const int dblocks = (int)((n + 3) / 4);
    lds = (size_t)256 * (c->L.depth + 2) * sizeof(int);
    hipLaunchKernelGGL(k_classify3_deep, dim3(dblocks), dim3(256), lds, c->stream, c->view, ev, g,
                       dv, plist, sidx, seed, mra, c->loc.as<uint8_t>(), c->starts.as<StartInfo>(),
                       c->amb.as<int32_t>(), c->bvalue.as<double>(), dq);
    return hipGetLastError();
}

int classify_parts(wos_ctx *c, const std::vector<Part> &parts, const EqView &ev,
                   const wos_run_params *prm, std::vector<double> &part_tol) {
    const int dim = c->mesh.dim;
    part_tol.assign(parts.size(), 0.0);
    {
        int64_t off = 0, poff = 0;
        for (size_t pi = 0; pi < parts.size(); ++pi) {
            const Part &p = parts[pi];
            const DomainInfo &D = c->L.domains[p.domain];
            const double *gb = D.bounds;
            const double scale = std::max({1.0, gb[1] - gb[0], gb[3] - gb[2], gb[5] - gb[4]});
            part_tol[pi] = 1e-10 * scale;
            GridSpec g{};
            if (p.grid) {
                g = *p.grid;
                for (int k = 0; k < 6; ++k) g.b[k] = gb[k];
            }
            DomainView dv{D.node_off, D.root_ref, (int)D.prim_off, (int)(D.prim_off + D.n_prims),
                          part_tol[pi], off, fp32_margin(D), p.domain};
            const double *dp = p.points ? c->plist_in.as<double>() + poff * dim : nullptr;
            const int64_t *ds = p.seed_index ? c->sidx.as<int64_t>() + poff : nullptr;
            hipError_t le;
            if (dim == 2) {
                // S lanes a point (k_classify2s) where 256 traversal stacks fit the LDS
                if ((size_t)256 * (c->L.depth + 2) * sizeof(int) <= c->lds_per_block)
                    le = launch_classify(c, k_classify2s<kClassify2Split>, ev, g, dv, dp, ds,
                                         p.count, prm->seed, prm->max_ray_attempts,
                                         256 / kClassify2Split, 256);
                else
                    le = launch_classify(c, k_classify<2>, ev, g, dv, dp, ds, p.count, prm->seed,
                                         prm->max_ray_attempts, kBlock, kBlock);
            } else {
                // the start query on 8 (or 4) subtrees where the block's
                // traversal stacks and result slots fit the LDS, else one
                // lane per point
                const size_t per_thread = (size_t)(c->L.depth + 2) * sizeof(int);
                const size_t fixed = 3 * 64 * sizeof(int);
                auto fits = [&](int split) {
                    return (size_t)(split + 3) * 64 * per_thread + fixed +
                               (size_t)split * 64 * sizeof(NpqD) <= c->lds_per_block;
                };
                if (fits(8))
                    le = launch_classify3<8>(c, ev, g, dv, dp, ds, p.count, prm->seed,
                                             prm->max_ray_attempts);
                else if (fits(4))
                    le = launch_classify3<4>(c, ev, g, dv, dp, ds, p.count, prm->seed,
                                             prm->max_ray_attempts);
                else
                    le = launch_classify(c, k_classify<3>, ev, g, dv, dp, ds, p.count, prm->seed,
                                         prm->max_ray_attempts, kBlock, kBlock);
            }
            HIPX(le);
            off += p.count;
            if (p.points) poff += p.count;
        }
    }
    return WOS_OK;
}

// Per-point result policy (runner.hpp:195-235) and the summary sums in
// sweep order (runner.hpp:224-228), for the whole solve and per part, from
// the host copies of the classification and the Welford state.
int finish_solve(wos_ctx *c, const std::vector<Part> &parts,
                 const std::vector<double> &part_tol, int64_t n,
                 const uint8_t *loc, const int32_t *amb, const double *bval,
                 const double *h_mean, const double *h_m2, const double *h_cnt,
                 const double *h_ss, const int32_t *h_hits, unsigned long long failed_walks,
                 double classify_s, double walk_s, double total_walk_s, int blocks,
                 const wos_run_params *prm, wos_point_stats *out,
                 wos_run_summary *summary, wos_run_summary *part_summaries) {
    wos_run_summary S{};
    int ii = 0;
    size_t pi = 0;
    int64_t part_end = parts.empty() ? 0 : parts[0].count;
    wos_run_summary P{};
    auto close_part = [&](size_t k) {
        const DomainInfo &D = c->L.domains[parts[k].domain];
        for (int j = 0; j < 6; ++j) P.grid_bounds[j] = D.bounds[j];
        P.boundary_tolerance = part_tol[k];
        if (part_summaries) part_summaries[k] = P;
        P = wos_run_summary{};
    };
    for (int64_t i = 0; i < n; ++i) {
        while (i >= part_end && pi + 1 < parts.size()) {
            close_part(pi);
            ++pi;
            part_end += parts[pi].count;
        }
        double mu = NAN, var = NAN, se = NAN, ms = NAN;
        int32_t hits = 0, walks = 0;
        if (loc[i] == 2) {
            mu = bval[i]; var = se = ms = 0.0;
        } else if (loc[i] == 1) {
            const double N = h_cnt[ii];
            mu = h_mean[ii];
            var = N > 1 ? h_m2[ii] / (N - 1.0) : 0.0;
            se = N > 0 ? std::sqrt(var / N) : 0.0;
            ms = N > 0 ? h_ss[ii] / N : 0.0;
            hits = h_hits[ii];
            walks = (int32_t)N;
            ++ii;
            for (wos_run_summary *T : {&S, &P}) {
                T->inside_count++;
                T->max_steps_hits += hits;
                T->sum_variance += var;
                T->sum_standard_error += se;
                T->sum_mean_steps += ms;
                T->total_walks += walks;
                T->total_steps += (int64_t)h_ss[ii - 1];
            }
        }
        if (loc[i] == 3) { S.indeterminate_points++; P.indeterminate_points++; }
        S.ambiguous_ray_retries += amb[i];
        P.ambiguous_ray_retries += amb[i];
        out->mean[i] = mu;
        if (out->variance) out->variance[i] = var;
        if (out->standard_error) out->standard_error[i] = se;
        if (out->mean_steps) out->mean_steps[i] = ms;
        if (out->location) out->location[i] = loc[i];
        if (out->max_steps_hits) out->max_steps_hits[i] = hits;
        if (out->walks) out->walks[i] = walks;
        if (out->ambiguous_ray_retries) out->ambiguous_ray_retries[i] = amb[i];
    }
    if (!parts.empty()) close_part(pi);
    S.wos_seconds = classify_s + total_walk_s;
    S.walk_kernel_seconds = walk_s;
    S.rounds = c->last_rounds;
    const DomainInfo &D0 = c->L.domains[parts[0].domain];
    for (int k = 0; k < 6; ++k) S.grid_bounds[k] = D0.bounds[k];
    S.boundary_tolerance = part_tol[0];
    S.launch_blocks = blocks;
    S.launch_threads = prm->precision == WOS_FP32_PERSISTENT ? c->last_threads : kBlock;
    S.lds_cached_nodes = prm->precision == WOS_FP32_PERSISTENT ? c->last_cached : 0;
    S.failed_walks = (int64_t)failed_walks;
    S.grid_build_seconds = c->grid_build_pending_s;
    c->grid_build_pending_s = 0.0;
    S.grid_bytes = c->grid_valid ? c->grid_bytes : 0;
    if (c->count_units) {
        unsigned long long u[7];
        HIPX(hipMemcpyAsync(u, c->units.p, sizeof(u), hipMemcpyDeviceToHost, c->stream));
        HIPX(hipStreamSynchronize(c->stream));
        S.q_queries = (int64_t)u[0]; S.q_grid_bound = (int64_t)u[1]; S.q_list_scans = (int64_t)u[2];
        S.q_list_entries = (int64_t)u[3]; S.q_list_prims = (int64_t)u[6];
        S.q_nodes = (int64_t)u[4]; S.q_leaf_prims = (int64_t)u[5];
        c->count_units = false;
    }
    if (summary) *summary = S;
    return WOS_OK;
}

int solve_once(wos_ctx *c, const std::vector<Part> &parts, const wos_eq_desc *eq,
               const wos_run_params *prm, wos_point_stats *out, wos_run_summary *summary,
               const AdjRequest *adjr = nullptr, wos_run_summary *part_summaries = nullptr) {
    if (!out || !out->mean) return fail(WOS_ERR_INPUT, "out->mean is required");
    EqView ev{};
    int64_t n = 0;
    int rc = prepare_solve(c, parts, eq, prm, adjr, ev, n);
    if (rc != WOS_OK) return rc;
    const int dim = c->mesh.dim;
    // the distance grid is built (once per mesh) outside the timed walks
    if (prm->precision == WOS_FP32_PERSISTENT) {
        HIPX(ensure_grid(c));
        HIPX(ensure_walk_stack(c));
    }
    HIPX(hipEventRecord(c->ev[0], c->stream));
    std::vector<double> part_tol;
    rc = classify_parts(c, parts, ev, prm, part_tol);
    if (rc != WOS_OK) return rc;
    HIPX(hipEventRecord(c->ev[1], c->stream));

    // classification results, then the statistics, reach the host through
    // page-locked staging (one DMA transfer each)
    HIPX(c->hcls.reserve((size_t)n * (sizeof(double) + sizeof(int32_t) + 1)));
    double *const bval = static_cast<double *>(c->hcls.p);
    int32_t *const amb = reinterpret_cast<int32_t *>(bval + n);
    uint8_t *const loc = reinterpret_cast<uint8_t *>(amb + n);
    HIPX(hipMemcpyAsync(loc, c->loc.p, n, hipMemcpyDeviceToHost, c->stream));
    HIPX(hipMemcpyAsync(amb, c->amb.p, n * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIPX(hipMemcpyAsync(bval, c->bvalue.p, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPX(hipStreamSynchronize(c->stream));

    std::vector<int32_t> inside;
    inside.reserve(n);
    for (int64_t i = 0; i < n; ++i)
        if (loc[i] == 1) inside.push_back((int32_t)i);
    const int n_in = (int)inside.size();

    c->last_rounds = 0;
    c->count_units = prm->count_units != 0 && prm->precision == WOS_FP32_PERSISTENT;
    if (c->count_units) {
        HIPX(c->units.reserve(7 * sizeof(unsigned long long)));
        hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, c->stream,
                           c->units.as<unsigned long long>(), 7);
        HIPX(hipGetLastError());
    }
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, c->stream,
                       c->failed.as<unsigned long long>(), 1);
    HIPX(hipGetLastError());
    HIPX(c->hstat.reserve((size_t)n_in * (4 * sizeof(double) + sizeof(int32_t))));
    double *const h_mean = static_cast<double *>(c->hstat.p);
    double *const h_m2 = h_mean + n_in;
    double *const h_cnt = h_m2 + n_in;
    double *const h_ss = h_cnt + n_in;
    int32_t *const h_hits = reinterpret_cast<int32_t *>(h_ss + n_in);
    double walk_s = 0.0, total_walk_s = 0.0;
    int blocks = 0;
    if (n_in > 0) {
        HIPX(upload(c->inside, inside, c->stream));
        HIPX(c->starts_f.reserve((size_t)n_in * sizeof(StartF)));
        hipLaunchKernelGGL(k_start_f, dim3((unsigned)((n_in + kBlock - 1) / kBlock)), dim3(kBlock),
                           0, c->stream, c->starts.as<StartInfo>(), c->inside.as<int32_t>(),
                           (int64_t)n_in, nullptr, c->starts_f.as<StartF>());
        HIPX(hipGetLastError());
        if (prm->precision == WOS_FP64_REFERENCE) {
            HIPX(c->ref_state.reserve((size_t)n_in * 8 * sizeof(uint64_t)));
            hipLaunchKernelGGL(k_ref_seed, dim3((unsigned)((n_in + kBlock - 1) / kBlock)),
                               dim3(kBlock), 0, c->stream, c->starts.as<StartInfo>(),
                               c->inside.as<int32_t>(), (int64_t)n_in, dim,
                               c->ref_state.as<uint64_t>());
            HIPX(hipGetLastError());
        }
        HIPX(c->mean.reserve(n_in * sizeof(double)));
        HIPX(c->m2.reserve(n_in * sizeof(double)));
        HIPX(c->count.reserve(n_in * sizeof(double)));
        HIPX(c->ssum.reserve(n_in * sizeof(double)));
        HIPX(c->hits.reserve(n_in * sizeof(int32_t)));
        HIPX(hipMemsetAsync(c->mean.p, 0, n_in * sizeof(double), c->stream));
        HIPX(hipMemsetAsync(c->m2.p, 0, n_in * sizeof(double), c->stream));
        HIPX(hipMemsetAsync(c->count.p, 0, n_in * sizeof(double), c->stream));
        HIPX(hipMemsetAsync(c->ssum.p, 0, n_in * sizeof(double), c->stream));
        HIPX(hipMemsetAsync(c->hits.p, 0, n_in * sizeof(int32_t), c->stream));
        HIPX(c->counter.reserve(kCounterBytes));
        const bool src = eq->source_kind != WOS_SRC_NONE;
        const bool scr = eq->kind == WOS_EQ_SCREENED || eq->kind == WOS_EQ_HELMHOLTZ;
        const bool adaptive = prm->target_se > 0.0;
        const int64_t cap = adaptive ? std::max<int64_t>(prm->max_walks, prm->n_walks) : prm->n_walks;
        // staging: one FP64 value + one int32 per walk, at most 4 GiB a
        // round; an adjoint replay adds 9 doubles + one int32 per walk and
        // may take up to 24 GiB (a quarter of the free device memory at
        // most), so that a solve's replay needs few rounds (each round ends
        // with a walker tail).  Results do not depend on the split.
        const int64_t walk_bytes = adjr ? 12 + 76 : 12;
        int64_t round_bytes = 4LL << 30;
        if (adjr) {
            size_t free_b = 0, total_b = 0;
            if (hipMemGetInfo(&free_b, &total_b) == hipSuccess)
                round_bytes = std::max(round_bytes, std::min<int64_t>((int64_t)(free_b / 4),
                                                                       24LL << 30));
        }
        const int64_t budget = std::max<int64_t>(
            std::min<int64_t>(round_bytes / walk_bytes, (1LL << 31) - 1), 1);
        AdjOut adj{};
        const bool want_v = adjr && adjr->grad_vertex && eq->boundary_kind == WOS_BND_VERTEX;
        const bool want_b = adjr && adjr->grad_params && eq->boundary_kind == WOS_BND_ID_CONST;
        if (adjr) {
            std::vector<double> coef(n_in);
            for (int k = 0; k < n_in; ++k) coef[k] = adjr->dl_du[inside[k]] / (double)prm->n_walks;
            HIPX(upload(c->coef, coef, c->stream));
            HIPX(c->scal_sum.reserve((size_t)n_in * 6 * sizeof(double)));
            HIPX(hipMemsetAsync(c->scal_sum.p, 0, (size_t)n_in * 6 * sizeof(double), c->stream));
            HIPX(c->grad_v.reserve(std::max<size_t>(1, c->mesh.n_verts()) * sizeof(double)));
            HIPX(hipMemsetAsync(c->grad_v.p, 0, c->mesh.n_verts() * sizeof(double), c->stream));
            HIPX(c->grad_b.reserve(WOS_MAX_BND_VALUES * sizeof(double)));
            HIPX(hipMemsetAsync(c->grad_b.p, 0, WOS_MAX_BND_VALUES * sizeof(double), c->stream));
        }
        std::vector<int64_t> done(n_in, 0), want(n_in, std::min<int64_t>(prm->n_walks, cap));
        // the FP32 merge keeps its lane states between rounds where a point's
        // walks may come in more than one
        double *lane_mu = nullptr, *lane_m2 = nullptr;
        if (prm->precision == WOS_FP32_PERSISTENT &&
            (adaptive || (int64_t)n_in * std::min<int64_t>(prm->n_walks, cap) > budget)) {
            HIPX(c->lane_st.reserve((size_t)n_in * 128 * sizeof(double)));
            lane_mu = c->lane_st.as<double>();
            lane_m2 = lane_mu + (size_t)n_in * 64;
        }
        // the SE test reads the counts and M2 into the statistics' staging
        double *const a_cnt = h_cnt, *const a_m2 = h_m2;
        std::vector<int32_t> all_ids(n_in);
        for (int k = 0; k < n_in; ++k) all_ids[k] = k;
        HIPX(upload(c->pid_all, all_ids, c->stream));
        // Rounds are queued one after the other without waiting: the host
        // plans the next round while the last one runs, and waits only where
        // it needs their results (the SE test of an adaptive solve, the end).
        int queued = 0;
        auto settle = [&]() -> hipError_t {
            if (queued == 0) return hipSuccess;
            const hipError_t e = hipEventSynchronize(c->round_ev[3 * (queued - 1) + 2]);
            if (e != hipSuccess) return e;
            for (int r = 0; r < queued; ++r) {
                walk_s += elapsed(c->round_ev[3 * r], c->round_ev[3 * r + 1]);
                total_walk_s += elapsed(c->round_ev[3 * r], c->round_ev[3 * r + 2]);
            }
            queued = 0;
            return hipSuccess;
        };
        int hslot = 0;
        for (int round = 0; round < 4096; ++round) {
            // ---- this round: uniform when every point needs the same walks
            int64_t remaining = 0;
            bool uniform = true;
            for (int k = 0; k < n_in; ++k) {
                remaining += want[k];
                if (want[k] != want[0] || done[k] != done[0]) uniform = false;
            }
            if (remaining == 0) {
                if (!adaptive) break;
                // SE test on every point below the cap (wos_run statistics)
                HIPX(settle());
                HIPX(hipMemcpyAsync(a_cnt, c->count.p, n_in * sizeof(double), hipMemcpyDeviceToHost, c->stream));
                HIPX(hipMemcpyAsync(a_m2, c->m2.p, n_in * sizeof(double), hipMemcpyDeviceToHost, c->stream));
                HIPX(hipStreamSynchronize(c->stream));
                const double t2 = prm->target_se * prm->target_se;
                for (int k = 0; k < n_in; ++k) {
                    const double N = a_cnt[k];
                    if (done[k] >= cap || N < 2) continue;
                    const double var = a_m2[k] / (N - 1.0);
                    if (std::sqrt(var / N) <= prm->target_se) continue;
                    const double need = std::ceil(1.1 * var / t2) - N;
                    want[k] = std::min<int64_t>(cap - done[k],
                                                std::max<int64_t>((int64_t)need, prm->n_walks));
                    remaining += want[k];
                }
                if (remaining == 0) break;
                uniform = false;
            }
            RoundView rv{};
            rv.starts = c->starts.as<StartInfo>();
            rv.sf = c->starts_f.as<StartF>();
            rv.inside = c->inside.as<int32_t>();
            if (uniform) {
                const int64_t cw = std::min<int64_t>(want[0], std::max<int64_t>(1, budget / n_in));
                rv.pid = c->pid_all.as<int32_t>();
                rv.pid_identity = 1;
                rv.n = n_in;
                set_walks_per_point(rv, (int)cw);
                rv.w0 = (int)done[0];
                rv.total = cw * n_in;
                for (int k = 0; k < n_in; ++k) { done[k] += cw; want[k] -= cw; }
            } else {
                // the round's points in sweep order until the budget is used
                // up, their tables written to the pinned buffer whose last
                // upload is done
                PinnedBuf &hb = c->h_round[hslot];
                if (c->h_round_used[hslot]) HIPX(hipEventSynchronize(c->h_round_ev[hslot]));
                const size_t cap_n = (size_t)n_in;
                HIPX(hb.reserve((cap_n + 1) * sizeof(int64_t) + 2 * cap_n * sizeof(int32_t)));
                int64_t *h_off = static_cast<int64_t *>(hb.p);
                int32_t *h_pid = reinterpret_cast<int32_t *>(h_off + cap_n + 1);
                int32_t *h_wbeg = h_pid + cap_n;
                int np = 0;
                int64_t tot = 0;
                h_off[0] = 0;
                for (int k = 0; k < n_in && tot < budget; ++k) {
                    if (want[k] == 0) continue;
                    const int64_t cnt = std::min<int64_t>(want[k], budget - tot);
                    h_pid[np] = k;
                    h_wbeg[np] = (int32_t)done[k];
                    tot += cnt;
                    h_off[++np] = tot;
                    done[k] += cnt;
                    want[k] -= cnt;
                }
                HIPX(c->r_pid.reserve(cap_n * sizeof(int32_t)));
                HIPX(c->r_wbeg.reserve(cap_n * sizeof(int32_t)));
                HIPX(c->r_off.reserve((cap_n + 1) * sizeof(int64_t)));
                HIPX(hipMemcpyAsync(c->r_pid.p, h_pid, (size_t)np * sizeof(int32_t),
                                    hipMemcpyHostToDevice, c->stream));
                HIPX(hipMemcpyAsync(c->r_wbeg.p, h_wbeg, (size_t)np * sizeof(int32_t),
                                    hipMemcpyHostToDevice, c->stream));
                HIPX(hipMemcpyAsync(c->r_off.p, h_off, (size_t)(np + 1) * sizeof(int64_t),
                                    hipMemcpyHostToDevice, c->stream));
                HIPX(hipEventRecord(c->h_round_ev[hslot], c->stream));
                c->h_round_used[hslot] = true;
                hslot ^= 1;
                // the point holding the first item of every block of 64 items
                const int64_t nb = (tot + (1 << kItemBlockShift) - 1) >> kItemBlockShift;
                HIPX(c->r_kblk.reserve((size_t)nb * sizeof(int32_t)));
                hipLaunchKernelGGL(k_round_kblk, dim3((unsigned)((nb + kBlock - 1) / kBlock)),
                                   dim3(kBlock), 0, c->stream, c->r_off.as<int64_t>(), np, nb,
                                   c->r_kblk.as<int32_t>());
                HIPX(hipGetLastError());
                rv.pid = c->r_pid.as<int32_t>();
                rv.wbeg = c->r_wbeg.as<int32_t>();
                rv.off = c->r_off.as<int64_t>();
                rv.kblk = c->r_kblk.as<int32_t>();
                rv.n = np;
                rv.total = tot;
            }
            HIPX(c->vals.reserve((size_t)rv.total * sizeof(double)));
            HIPX(c->steps.reserve((size_t)rv.total * sizeof(int32_t)));
            HIPX(reserve_terminal(c, ev, prm->precision, adjr != nullptr, rv.total));
            if (adjr) {
                HIPX(c->adj_scal.reserve((size_t)rv.total * 6 * sizeof(double)));
                HIPX(c->adj_vw.reserve((size_t)rv.total * 3 * sizeof(double)));
                HIPX(c->adj_prim.reserve((size_t)rv.total * sizeof(int32_t)));
                adj.scal = c->adj_scal.as<double>();
                adj.vw = c->adj_vw.as<double>();
                adj.prim = c->adj_prim.as<int32_t>();
            }
            // at most 64 rounds queued (the timing events they hold)
            if (queued == 64) HIPX(settle());
            while ((int)c->round_ev.size() < 3 * (queued + 1)) {
                hipEvent_t e = nullptr;
                HIPX(hipEventCreate(&e));
                c->round_ev.push_back(e);
            }
            hipEvent_t *const re = &c->round_ev[3 * queued];
            HIPX(hipEventRecord(re[0], c->stream));
            HIPX(dispatch_walk(c, dim, adjr != nullptr, ev, src, scr, prm->precision, rv,
                               prm->epsilon, prm->max_steps, adj, &blocks));
            HIPX(hipEventRecord(re[1], c->stream));
            const bool term = prm->precision == WOS_FP32_PERSISTENT && !adjr &&
                              ev.boundary_kind != WOS_BND_ZERO;
            HIPX(launch_merge(c->stream, prm->precision, rv, rv.n, c->vals.as<double>(),
                              c->steps.as<int32_t>(), c->mean.as<double>(), c->m2.as<double>(),
                              c->count.as<double>(), c->ssum.as<double>(), c->hits.as<int32_t>(),
                              c->failed.as<unsigned long long>(), lane_mu, lane_m2, dim, c->view,
                              ev, term ? c->term_pw.as<float4>() : nullptr,
                              term ? c->term_slot.as<int32_t>() : nullptr));
            if (adjr) {
                hipLaunchKernelGGL(k_adj_reduce, dim3((rv.n + kBlock - 1) / kBlock), dim3(kBlock), 0,
                                   c->stream, rv, adj, c->view, c->coef.as<double>(),
                                   c->scal_sum.as<double>(), want_v ? c->grad_v.as<double>() : nullptr,
                                   want_b ? c->grad_b.as<double>() : nullptr);
                HIPX(hipGetLastError());
            }
            HIPX(hipEventRecord(re[2], c->stream));
            ++queued;
            ++c->last_rounds;
        }
        HIPX(settle());
        HIPX(hipMemcpyAsync(h_mean, c->mean.p, n_in * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPX(hipMemcpyAsync(h_m2, c->m2.p, n_in * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPX(hipMemcpyAsync(h_cnt, c->count.p, n_in * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPX(hipMemcpyAsync(h_ss, c->ssum.p, n_in * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPX(hipMemcpyAsync(h_hits, c->hits.p, n_in * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
        HIPX(hipStreamSynchronize(c->stream));
    }
    if (adjr) {
        // boundary points (u = g at the start's nearest point), then the
        // parameter gradients summed over points in sweep order
        const bool want_v = adjr->grad_vertex && eq->boundary_kind == WOS_BND_VERTEX;
        const bool want_b = adjr->grad_params && eq->boundary_kind == WOS_BND_ID_CONST;
        if (n_in == 0) {
            HIPX(c->grad_v.reserve(std::max<size_t>(1, c->mesh.n_verts()) * sizeof(double)));
            HIPX(hipMemsetAsync(c->grad_v.p, 0, c->mesh.n_verts() * sizeof(double), c->stream));
            HIPX(c->grad_b.reserve(WOS_MAX_BND_VALUES * sizeof(double)));
            HIPX(hipMemsetAsync(c->grad_b.p, 0, WOS_MAX_BND_VALUES * sizeof(double), c->stream));
        }
        std::vector<double> dl(adjr->dl_du, adjr->dl_du + n);
        HIPX(upload(c->dl_du, dl, c->stream));
        hipLaunchKernelGGL(k_adj_boundary, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                           c->stream, c->view, c->loc.as<uint8_t>(), c->starts.as<StartInfo>(),
                           c->dl_du.as<double>(), n, want_v ? c->grad_v.as<double>() : nullptr,
                           want_b ? c->grad_b.as<double>() : nullptr);
        HIPX(hipGetLastError());
        std::vector<double> scal((size_t)n_in * 6), gb(WOS_MAX_BND_VALUES, 0.0);
        if (n_in > 0)
            HIPX(hipMemcpyAsync(scal.data(), c->scal_sum.p, scal.size() * sizeof(double),
                                hipMemcpyDeviceToHost, c->stream));
        HIPX(hipMemcpyAsync(gb.data(), c->grad_b.p, WOS_MAX_BND_VALUES * sizeof(double),
                            hipMemcpyDeviceToHost, c->stream));
        if (adjr->grad_vertex) {
            if (want_v)
                HIPX(hipMemcpyAsync(adjr->grad_vertex, c->grad_v.p, c->mesh.n_verts() * sizeof(double),
                                    hipMemcpyDeviceToHost, c->stream));
            else
                std::fill(adjr->grad_vertex, adjr->grad_vertex + c->mesh.n_verts(), 0.0);
        }
        HIPX(hipStreamSynchronize(c->stream));
        if (adjr->grad_params) {
            double g6[6] = {0, 0, 0, 0, 0, 0};
            for (int k = 0; k < n_in; ++k) {
                const double cf = adjr->dl_du[inside[k]] / (double)prm->n_walks;
                for (int i = 0; i < 6; ++i) g6[i] += cf * scal[(size_t)k * 6 + i];
            }
            double *gp = adjr->grad_params;
            // alpha: pathwise only where the walk's draws do not depend on it
            const bool alpha_ok = eq->kind == WOS_EQ_SCREENED && dim == 3 &&
                                  (eq->source_kind == WOS_SRC_NONE || eq->source_mode == WOS_SOURCE_UNIFORM);
            gp[WOS_GRAD_ALPHA] = eq->kind == WOS_EQ_HELMHOLTZ ? NAN
                                 : eq->kind != WOS_EQ_SCREENED ? 0.0 : (alpha_ok ? g6[0] : NAN);
            gp[WOS_GRAD_SRC_VALUE] = g6[1];
            gp[WOS_GRAD_SRC_RADIUS] = eq->source_kind == WOS_SRC_GAUSSIAN ? g6[2] : 0.0;
            for (int i = 0; i < 3; ++i)
                gp[WOS_GRAD_SRC_CENTER + i] = eq->source_kind == WOS_SRC_GAUSSIAN ? g6[3 + i] : 0.0;
            for (int i = 0; i < WOS_MAX_BND_VALUES; ++i)
                gp[WOS_GRAD_BND_VALUES + i] = want_b ? gb[i] : 0.0;
        }
    }
    const double classify_s = elapsed(c->ev[0], c->ev[1]);

    // per-point result policy (runner.hpp:195-235) and the summary sums in
    // sweep order (runner.hpp:224-228), for the whole solve and per part
    unsigned long long nf = 0;
    HIPX(hipMemcpyAsync(&nf, c->failed.p, sizeof(nf), hipMemcpyDeviceToHost, c->stream));
    HIPX(hipStreamSynchronize(c->stream));
    return finish_solve(c, parts, part_tol, n, loc, amb, bval, h_mean, h_m2, h_cnt, h_ss, h_hits,
                        nf, classify_s,
                        walk_s, total_walk_s, blocks, prm, out, summary, part_summaries);
}

// ---------------------------------------------------------------------------
// graph replay (params->use_graph = 1): a fixed-walk FP32 grid solve whose
// whole device sequence -- classification, inside-point compaction, Welford
// reset, persistent walker, Welford merge -- is captured once into a
// hipGraph and replayed while the problem, the tuning and every buffer it
// touches stay the same (any change is detected and captured again).  No
// memset nodes: buffers are cleared by kernels.  Results equal solve_once's
// bit for bit.
// ---------------------------------------------------------------------------
template <typename T>
void key_put(std::vector<uint8_t> &k, const T &v) {
    const uint8_t *b = reinterpret_cast<const uint8_t *>(&v);
    k.insert(k.end(), b, b + sizeof(T));
}

void drop_graph(wos_ctx *c) {
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->graph_exec = nullptr;
    c->graph = nullptr;
    c->graph_key.clear();
}

int solve_graph(wos_ctx *c, const std::vector<Part> &parts, const wos_eq_desc *eq,
                const wos_run_params *prm, wos_point_stats *out, wos_run_summary *summary,
                wos_run_summary *part_summaries, bool capture = true) {
    if (!out || !out->mean) return fail(WOS_ERR_INPUT, "out->mean is required");
    if (!prm) return fail(WOS_ERR_INPUT, "params are required");
    if (prm->precision != WOS_FP32_PERSISTENT || prm->target_se > 0.0)
        return fail(WOS_ERR_INPUT, "graph replay runs fixed-walk FP32 persistent solves "
                                   "(precision FP32, target_se = 0)");
    for (const Part &p : parts)
        if (!p.grid) return fail(WOS_ERR_INPUT, "graph replay runs grid solves");
    EqView ev{};
    int64_t n = 0;
    int rc = prepare_solve(c, parts, eq, prm, nullptr, ev, n);
    if (rc != WOS_OK) return rc;
    const int64_t nw = prm->n_walks;
    if (n * nw > ((int64_t)4 << 30) / 12)
        return fail(WOS_ERR_INPUT, "graph replay needs points x n_walks <= 357913941 (one round)");
    const int dim = c->mesh.dim;
    const int64_t nb = (n + kBlock - 1) / kBlock;
    HIPX(c->inside.reserve((size_t)n * sizeof(int32_t)));
    HIPX(c->starts_f.reserve((size_t)n * sizeof(StartF)));
    HIPX(c->g_bcount.reserve((size_t)nb * sizeof(int32_t)));
    HIPX(c->g_boff.reserve((size_t)nb * sizeof(int64_t)));
    HIPX(c->g_misc.reserve(64));
    HIPX(c->mean.reserve((size_t)n * sizeof(double)));
    HIPX(c->m2.reserve((size_t)n * sizeof(double)));
    HIPX(c->count.reserve((size_t)n * sizeof(double)));
    HIPX(c->ssum.reserve((size_t)n * sizeof(double)));
    HIPX(c->hits.reserve((size_t)n * sizeof(int32_t)));
    HIPX(c->vals.reserve((size_t)(n * nw) * sizeof(double)));
    HIPX(c->steps.reserve((size_t)(n * nw) * sizeof(int32_t)));
    HIPX(reserve_terminal(c, ev, WOS_FP32_PERSISTENT, false, n * nw));
    HIPX(c->counter.reserve(kCounterBytes));
    if (c->pid_all.bytes < (size_t)n * sizeof(int32_t)) {
        std::vector<int32_t> ids((size_t)n);
        for (int64_t k = 0; k < n; ++k) ids[(size_t)k] = (int32_t)k;
        HIPX(upload(c->pid_all, ids, c->stream));
    }
    const PackLayout PL(n);
    HIPX(c->pack.reserve(PL.total));
    HIPX(c->hpack.reserve(PL.total));
    HIPX(ensure_grid(c));
    HIPX(ensure_walk_stack(c));
    c->count_units = false;
    int32_t *n_in_dev = c->g_misc.as<int32_t>();
    int64_t *total_dev = reinterpret_cast<int64_t *>(c->g_misc.as<uint8_t>() + 8);

    std::vector<uint8_t> key;
    key_put(key, n);
    for (const Part &p : parts) { key_put(key, p.domain); key_put(key, *p.grid); }
    key_put(key, ev);
    key_put(key, *prm);
    key_put(key, c->tune);
    key_put(key, c->view);
    for (const DevBuf *b : {&c->loc, &c->starts, &c->starts_f, &c->amb, &c->bvalue, &c->inside, &c->g_bcount,
                            &c->g_boff, &c->g_misc, &c->mean, &c->m2, &c->count, &c->ssum,
                            &c->hits, &c->vals, &c->steps, &c->counter, &c->pid_all, &c->vv,
                            &c->failed, &c->wstack, &c->pack,
                            &c->field, &c->term_pw, &c->term_slot})
        key_put(key, b->p);
    auto body = [&]() -> int {
            int r = classify_parts(c, parts, ev, prm, c->graph_tol);
            if (r != WOS_OK) return r;
            hipLaunchKernelGGL(k_inside_count, dim3((unsigned)nb), dim3(kBlock), 0, c->stream,
                               c->loc.as<uint8_t>(), n, c->g_bcount.as<int32_t>());
            hipLaunchKernelGGL(k_inside_scan, dim3(1), dim3(1024), 0, c->stream,
                               c->g_bcount.as<int32_t>(), nb, c->g_boff.as<int64_t>(), n_in_dev,
                               total_dev, nw);
            hipLaunchKernelGGL(k_inside_scatter, dim3((unsigned)nb), dim3(kBlock), 0, c->stream,
                               c->loc.as<uint8_t>(), n, c->g_boff.as<int64_t>(),
                               c->inside.as<int32_t>(), c->starts.as<StartInfo>(),
                               c->starts_f.as<StartF>());
            hipLaunchKernelGGL(k_zero_stats, dim3(256), dim3(kBlock), 0, c->stream,
                               c->mean.as<double>(), c->m2.as<double>(), c->count.as<double>(),
                               c->ssum.as<double>(), c->hits.as<int32_t>(),
                               c->failed.as<unsigned long long>(), n);
            HIPX(hipGetLastError());
            RoundView rv{};
            rv.starts = c->starts.as<StartInfo>();
            rv.sf = c->starts_f.as<StartF>();
            rv.inside = c->inside.as<int32_t>();
            rv.pid = c->pid_all.as<int32_t>();
            rv.pid_identity = 1;
            rv.n = (int)n;
            set_walks_per_point(rv, (int)nw);
            rv.w0 = 0;
            rv.total = n * nw;
            rv.n_dev = n_in_dev;
            rv.total_dev = total_dev;
            if (!capture) HIPX(hipEventRecord(c->ev[2], c->stream));
            HIPX(dispatch_walk(c, dim, false, ev, eq->source_kind != WOS_SRC_NONE,
                               eq->kind == WOS_EQ_SCREENED || eq->kind == WOS_EQ_HELMHOLTZ,
                               WOS_FP32_PERSISTENT, rv, prm->epsilon,
                               prm->max_steps, AdjOut{}, &c->graph_blocks));
            if (!capture) HIPX(hipEventRecord(c->ev[3], c->stream));
            const bool term = ev.boundary_kind != WOS_BND_ZERO;
            HIPX(launch_merge(c->stream, WOS_FP32_PERSISTENT, rv, n, c->vals.as<double>(),
                              c->steps.as<int32_t>(), c->mean.as<double>(), c->m2.as<double>(),
                              c->count.as<double>(), c->ssum.as<double>(), c->hits.as<int32_t>(),
                              c->failed.as<unsigned long long>(), nullptr, nullptr, dim, c->view,
                              ev, term ? c->term_pw.as<float4>() : nullptr,
                              term ? c->term_slot.as<int32_t>() : nullptr));
            hipLaunchKernelGGL(k_pack_results, dim3((unsigned)nb), dim3(kBlock), 0, c->stream, n,
                               n_in_dev, c->failed.as<unsigned long long>(), c->loc.as<uint8_t>(),
                               c->amb.as<int32_t>(), c->bvalue.as<double>(), c->mean.as<double>(),
                               c->m2.as<double>(), c->count.as<double>(), c->ssum.as<double>(),
                               c->hits.as<int32_t>(), c->pack.as<unsigned char>());
            HIPX(hipGetLastError());
            return WOS_OK;
        };
    double classify_s = 0.0, walk_s, total_walk_s;
    if (!capture) {
        // the same device sequence launched directly: no host round trip
        // between classification and walks
        HIPX(hipEventRecord(c->ev[0], c->stream));
        const int brc = body();
        if (brc != WOS_OK) return brc;
        HIPX(hipEventRecord(c->ev[1], c->stream));
        HIPX(hipStreamSynchronize(c->stream));
        classify_s = elapsed(c->ev[0], c->ev[2]);
        walk_s = elapsed(c->ev[2], c->ev[3]);
        total_walk_s = elapsed(c->ev[2], c->ev[1]);
    } else {
        if (key != c->graph_key) {
            drop_graph(c);
            HIPX(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
            const int brc = body();
            hipGraph_t g = nullptr;
            const hipError_t ec = hipStreamEndCapture(c->stream, &g);
            if (brc != WOS_OK) { if (g) (void)hipGraphDestroy(g); return brc; }
            HIPX(ec);
            c->graph = g;
            HIPX(hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0));
            c->graph_key = key;
        }
        HIPX(hipEventRecord(c->ev[0], c->stream));
        HIPX(hipGraphLaunch(c->graph_exec, c->stream));
        HIPX(hipEventRecord(c->ev[1], c->stream));
        HIPX(hipStreamSynchronize(c->stream));
        // the graph runs classification and walks back to back: its time is
        // reported as both the solve and the walker time
        walk_s = total_walk_s = elapsed(c->ev[0], c->ev[1]);
    }

    // every result in one transfer into page-locked memory
    HIPX(hipMemcpyAsync(c->hpack.p, c->pack.p, PL.total, hipMemcpyDeviceToHost, c->stream));
    HIPX(hipStreamSynchronize(c->stream));
    const unsigned char *P = static_cast<const unsigned char *>(c->hpack.p);
    const unsigned long long n_failed = reinterpret_cast<const unsigned long long *>(P)[1];
    c->last_rounds = 1;
    return finish_solve(c, parts, c->graph_tol, n, P + PL.loc,
                        reinterpret_cast<const int32_t *>(P + PL.amb),
                        reinterpret_cast<const double *>(P + PL.bval),
                        reinterpret_cast<const double *>(P + PL.mean),
                        reinterpret_cast<const double *>(P + PL.m2),
                        reinterpret_cast<const double *>(P + PL.cnt),
                        reinterpret_cast<const double *>(P + PL.ss),
                        reinterpret_cast<const int32_t *>(P + PL.hits), n_failed, classify_s,
                        walk_s, total_walk_s, c->graph_blocks, prm, out, summary, part_summaries);
}

// ---------------------------------------------------------------------------
// autotuner.  It searches launch-shape fields that never change a result
// (walkers per block, lanes per walk-step phase, refill chunk, LDS node
// cache, xcd_parts, the LDS share of the traversal stack): every walk's
// draws and queries are independent of them, so a tuned solve is
// bit-identical to an untuned one.  With params->tune = 2 it also searches
// the BVH width, which can change the order of a relaxed tree search (and
// so a walk) where the grid has no list, not the estimator's expectation.
// The distance grid, its list cap and relax stay as set.  Coordinate descent from the current
// shape over a fixed candidate list, each candidate timed twice (walker
// kernel time, the smaller counts) on a pilot solve of the same problem with
// about 16 M walks.
// ---------------------------------------------------------------------------
TuneKey tune_key(const wos_ctx *c, const std::vector<Part> &parts, const wos_eq_desc *eq,
                 int mode) {
    TuneKey k{};
    k.mode = mode;
    k.dim = c->mesh.dim;
    k.n_domains = (int)c->L.domains.size();
    k.n_prims = c->view.n_prims;
    for (const Part &p : parts) k.n_points += p.count;
    k.kind = eq ? eq->kind : -1;
    k.source_mode = eq ? eq->source_mode : -1;
    k.source_kind = eq ? eq->source_kind : -1;
    k.grid_cells = c->tune.grid_cells;
    return k;
}

int autotune(wos_ctx *c, const std::vector<Part> &parts, const wos_eq_desc *eq,
             const wos_run_params *prm, wos_tuning *chosen) {
    int64_t n = 0;
    for (const Part &p : parts) n += p.count;
    if (n <= 0) return fail(WOS_ERR_INPUT, "point count out of range");
    wos_run_params p = *prm;
    p.tune = 0;
    p.target_se = 0.0;
    p.count_units = 0;
    p.use_graph = 0;
    p.precision = WOS_FP32_PERSISTENT;
    const int64_t want = ((int64_t)1 << 24) / n + 1;
    p.n_walks = (int)std::max<int64_t>(16, std::min<int64_t>(want, std::max(16, prm->n_walks)));
    std::vector<double> mean((size_t)n);
    wos_point_stats st{};
    st.mean = mean.data();
    const wos_tuning saved = c->tune;
    auto timed = [&](const wos_tuning &t, double *sec) -> int {
        c->tune = t;
        double best = INFINITY;
        for (int r = 0; r < 2; ++r) {
            wos_run_summary S{};
            const int rc = solve_once(c, parts, eq, &p, &st, &S);
            if (rc != WOS_OK) return rc;
            best = std::min(best, S.walk_kernel_seconds);
        }
        *sec = best;
        return WOS_OK;
    };
    wos_tuning cur = saved;
    double cur_s = 0.0;
    int rc = timed(cur, &cur_s);
    if (rc != WOS_OK) { c->tune = saved; return rc; }
    static const int kReady[] = {16, 24, 32, 40, 48, 56};
    static const int kThreads[] = {128, 192, 256};
    static const int kChunk[] = {32, 64, 256};
    static const int kLds[] = {0, -1};
    static const int kParts[] = {1, 8};
    static const int kStack[] = {0, 8, 32};
    static const int kWidth[] = {2, 4};
    auto descend = [&](int wos_tuning::*field, const int *vals, int nv) -> int {
        for (int i = 0; i < nv; ++i) {
            if (cur.*field == vals[i]) continue;
            wos_tuning t = cur;
            t.*field = vals[i];
            double sec = 0.0;
            const int r = timed(t, &sec);
            if (r != WOS_OK) return r;
            if (sec < cur_s) { cur = t; cur_s = sec; }
        }
        return WOS_OK;
    };
    if ((rc = descend(&wos_tuning::min_ready, kReady, 6)) == WOS_OK &&
        (rc = descend(&wos_tuning::block_threads, kThreads, 3)) == WOS_OK &&
        (rc = descend(&wos_tuning::chunk, kChunk, 3)) == WOS_OK &&
        (rc = descend(&wos_tuning::lds_nodes, kLds, 2)) == WOS_OK &&
        (rc = descend(&wos_tuning::xcd_parts, kParts, 2)) == WOS_OK &&
        (rc = descend(&wos_tuning::stack_lds, kStack, 3)) == WOS_OK && prm->tune == 2)
        rc = descend(&wos_tuning::bvh_width, kWidth, 2);
    c->tune = saved;
    if (rc != WOS_OK) return rc;
    *chosen = cur;
    c->tuned.emplace_back(tune_key(c, parts, eq, prm->tune == 2 ? 2 : 1), cur);
    return WOS_OK;
}

// solve_once with the autotuned launch shape when params->tune = 1 (FP32):
// the shape cached for this problem, found by autotune() on first use.
// A fixed-walk FP32 forward grid solve that fits one round runs the graph's
// device sequence directly (device-side compaction of the inside points);
// everything else takes the general path of solve_once.
bool device_compaction_ok(const wos_ctx *c, const std::vector<Part> &parts,
                          const wos_run_params *prm, const AdjRequest *adjr) {
    if (!c || !c->has_mesh || !prm || adjr || prm->precision != WOS_FP32_PERSISTENT ||
        prm->target_se != 0.0 || prm->count_units || prm->n_walks < 1)
        return false;
    int64_t n = 0;
    for (const Part &p : parts) {
        if (!p.grid) return false;
        n += p.count;
    }
    return n > 0 && n * (int64_t)prm->n_walks <= ((int64_t)4 << 30) / 12;
}

int solve(wos_ctx *c, const std::vector<Part> &parts, const wos_eq_desc *eq,
          const wos_run_params *prm, wos_point_stats *out, wos_run_summary *summary,
          const AdjRequest *adjr = nullptr, wos_run_summary *part_summaries = nullptr) {
    if (c && prm && prm->use_graph) {
        if (adjr) return fail(WOS_ERR_INPUT, "graph replay runs forward solves only");
        if (prm->use_graph != 1) return fail(WOS_ERR_INPUT, "use_graph must be 0 or 1");
    }
    auto run = [&]() {
        if (prm && prm->use_graph) return solve_graph(c, parts, eq, prm, out, summary, part_summaries);
        if (device_compaction_ok(c, parts, prm, adjr))
            return solve_graph(c, parts, eq, prm, out, summary, part_summaries, false);
        return solve_once(c, parts, eq, prm, out, summary, adjr, part_summaries);
    };
    if (!c || !prm || prm->tune == 0 || prm->precision != WOS_FP32_PERSISTENT || !c->has_mesh)
        return run();
    if (prm->tune != 1 && prm->tune != 2) return fail(WOS_ERR_INPUT, "tune must be 0, 1 or 2");
    const TuneKey key = tune_key(c, parts, eq, prm->tune);
    const wos_tuning *found = nullptr;
    for (const auto &e : c->tuned)
        if (e.first == key) found = &e.second;
    wos_tuning shape;
    if (found) {
        shape = *found;
    } else {
        const int rc = autotune(c, parts, eq, prm, &shape);
        if (rc != WOS_OK) return rc;
    }
    const wos_tuning saved = c->tune;
    c->tune.block_threads = shape.block_threads;
    c->tune.min_ready = shape.min_ready;
    c->tune.chunk = shape.chunk;
    c->tune.lds_nodes = shape.lds_nodes;
    c->tune.xcd_parts = shape.xcd_parts;
    c->tune.stack_lds = shape.stack_lds;
    if (prm->tune == 2) c->tune.bvh_width = shape.bvh_width;
    const int rc = run();
    c->tune = saved;
    return rc;
}

}  // namespace

extern "C" {

const char *wos_last_error(void) { return g_last_error.c_str(); }

const char *wos_version(void) { return "wos-mi355x 0.1 (gfx950)"; }

wos_ctx *wos_ctx_create(int device) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        fail(WOS_ERR_DEVICE, "no HIP device available");
        return nullptr;
    }
    if (device < 0 || device >= count) {
        fail(WOS_ERR_INPUT, "device index out of range");
        return nullptr;
    }
    auto c = std::make_unique<wos_ctx>();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
        fail(WOS_ERR_DEVICE, "could not create a HIP stream");
        return nullptr;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        fail(WOS_ERR_DEVICE, "hipGetDeviceProperties failed");
        return nullptr;
    }
    c->n_cu = prop.multiProcessorCount;
    c->lds_per_block = prop.sharedMemPerBlock;
    c->lds_per_cu = prop.maxSharedMemoryPerMultiProcessor > 0
                        ? prop.maxSharedMemoryPerMultiProcessor : prop.sharedMemPerBlock;
    for (auto &e : c->ev)
        if (hipEventCreate(&e) != hipSuccess) {
            fail(WOS_ERR_DEVICE, "hipEventCreate failed");
            return nullptr;
        }
    for (auto &e : c->h_round_ev)
        if (hipEventCreate(&e) != hipSuccess) {
            fail(WOS_ERR_DEVICE, "hipEventCreate failed");
            return nullptr;
        }
    {
        const hipError_t e = scratch_address(device, &c->scratch_dev);
        if (e != hipSuccess) {
            fail(WOS_ERR_DEVICE, std::string("resolving the scratch array: ") + hipGetErrorString(e));
            return nullptr;
        }
    }
    return c.release();
}

void wos_ctx_destroy(wos_ctx *c) { delete c; }

int wos_ctx_set_domains(wos_ctx *c, int dim, int n_domains, const double *const *verts,
                        const int64_t *n_verts, const int32_t *const *prims,
                        const int64_t *n_prims, const int32_t *const *boundary_ids) {
    if (!c) return fail(WOS_ERR_INPUT, "context is NULL");
    if (dim != 2 && dim != 3) return fail(WOS_ERR_INPUT, "dim must be 2 or 3");
    if (n_domains < 1 || !verts || !n_verts || !prims || !n_prims)
        return fail(WOS_ERR_INPUT, "need at least one domain");
    std::vector<HostMesh> meshes(n_domains);
    std::vector<HostBVH> bvhs(n_domains);
    HostMesh all;
    all.dim = dim;
    int64_t vo = 0, po = 0;
    for (int d = 0; d < n_domains; ++d) {
        if (!verts[d] || !prims[d] || n_verts[d] <= 0 || n_prims[d] <= 0)
            return fail(WOS_ERR_MESH, "mesh has no vertices or primitives");
        HostMesh &m = meshes[d];
        m.dim = dim;
        m.verts.assign(verts[d], verts[d] + n_verts[d] * dim);
        m.prims.assign(prims[d], prims[d] + n_prims[d] * dim);
        m.bids.assign(n_prims[d], 0);
        if (boundary_ids && boundary_ids[d] && dim == 2)
            m.bids.assign(boundary_ids[d], boundary_ids[d] + n_prims[d]);
        for (double v : m.verts)
            if (!std::isfinite(v)) return fail(WOS_ERR_MESH, "vertex coordinates must be finite");
        for (int32_t i : m.prims)
            if (i < 0 || i >= n_verts[d]) return fail(WOS_ERR_MESH, "vertex index out of range");
        std::string err;
        bvhs[d] = build_reference_bvh(m, &err);
        if (bvhs[d].nodes.empty()) return fail(WOS_ERR_MESH, err);
        if (bvhs[d].depth > 60) return fail(WOS_ERR_MESH, "BVH deeper than the 60-level traversal stack");
        // the concatenated mesh: global vertex and primitive numbering
        all.verts.insert(all.verts.end(), m.verts.begin(), m.verts.end());
        for (int32_t i : m.prims) all.prims.push_back((int32_t)(i + vo));
        all.bids.insert(all.bids.end(), m.bids.begin(), m.bids.end());
        vo += n_verts[d];
        po += n_prims[d];
    }
    if (po >= (1 << 27)) return fail(WOS_ERR_MESH, "more than 2^27 primitives in total");
    std::vector<const HostMesh *> mp;
    std::vector<const HostBVH *> bp;
    for (int d = 0; d < n_domains; ++d) { mp.push_back(&meshes[d]); bp.push_back(&bvhs[d]); }
    DeviceLayouts L = make_layouts(mp, bp);
    // From here on the old mesh's buffers are replaced in place: until every
    // upload has succeeded the context holds no mesh, so a failed upload
    // leaves it refusing solves instead of reading freed buffers.
    c->has_mesh = false;
    c->grid_valid = false;
    drop_graph(c);
    HIPX(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    if (dim == 2) {
        HIPX(upload(c->nd2, L.nd2, s));
        HIPX(upload(c->segd, L.segd, s));
        HIPX(upload(c->nf2, L.nf2, s));
        HIPX(upload(c->nw2, L.nw2, s));
        HIPX(upload(c->segf, L.segf, s));
    } else {
        HIPX(upload(c->nd3, L.nd3, s));
        HIPX(upload(c->trid, L.trid, s));
        HIPX(upload(c->nf3, L.nf3, s));
        HIPX(upload(c->nw3, L.nw3, s));
        HIPX(upload(c->trif, L.trif, s));
    }
    std::vector<DomainF> doms, doms_w;
    for (const DomainInfo &D : L.domains) {
        doms.push_back(DomainF{D.root_ref, fp32_margin(D)});
        doms_w.push_back(DomainF{D.root_ref_w, fp32_margin(D)});
    }
    HIPX(upload(c->doms, doms, s));
    HIPX(upload(c->doms_w, doms_w, s));
    HIPX(upload(c->slot_prim, L.slot_prim, s));
    HIPX(upload(c->prim_bid, all.bids, s));
    HIPX(upload(c->prims, all.prims, s));
    HIPX(upload(c->verts, all.verts, s));
    MeshView v{};
    v.dim = dim;
    v.n_prims = (int)po;
    v.n_nodes = dim == 2 ? (int)L.nd2.size() : (int)L.nd3.size();
    v.depth = L.depth;
    v.nd2 = c->nd2.as<NodeD2>(); v.nd3 = c->nd3.as<NodeD3>();
    v.segd = c->segd.as<SegD>(); v.trid = c->trid.as<TriD>();
    v.slot_prim = c->slot_prim.as<int32_t>(); v.prim_bid = c->prim_bid.as<int32_t>();
    v.prims = c->prims.as<int32_t>(); v.verts = c->verts.as<double>();
    v.nf2 = c->nf2.as<NodeF2>(); v.nf3 = c->nf3.as<NodeF3>();
    v.nw2 = c->nw2.as<NodeF2W>(); v.nw3 = c->nw3.as<NodeF3W>();
    v.segf = c->segf.as<SegF>(); v.trif = c->trif.as<TriF>();
    v.doms = c->doms.as<DomainF>();
    c->view = v;
    c->mesh = std::move(all);
    c->L = std::move(L);
    c->has_mesh = true;
    c->tuned.clear();
    return WOS_OK;
}

int wos_ctx_set_mesh(wos_ctx *c, int dim, const double *verts, int64_t n_verts,
                     const int32_t *prims, int64_t n_prims, const int32_t *boundary_ids) {
    if (!c) return fail(WOS_ERR_INPUT, "context is NULL");
    if (!verts || !prims || n_verts <= 0 || n_prims <= 0)
        return fail(WOS_ERR_MESH, "mesh has no vertices or primitives");
    return wos_ctx_set_domains(c, dim, 1, &verts, &n_verts, &prims, &n_prims, &boundary_ids);
}

int64_t wos_bvh_build_host(int dim, const double *verts, int64_t n_verts, const int32_t *prims,
                           int64_t n_prims, int32_t *count, int32_t *idx, double *lo, double *hi,
                           int32_t *slots) {
    if ((dim != 2 && dim != 3) || !verts || !prims || n_verts <= 0 || n_prims <= 0) {
        fail(WOS_ERR_INPUT, "invalid mesh arrays");
        return -1;
    }
    HostMesh m;
    m.dim = dim;
    m.verts.assign(verts, verts + n_verts * dim);
    m.prims.assign(prims, prims + n_prims * dim);
    for (int32_t i : m.prims)
        if (i < 0 || i >= n_verts) {
            fail(WOS_ERR_MESH, "vertex index out of range");
            return -1;
        }
    std::string err;
    HostBVH b = build_reference_bvh(m, &err);
    if (b.nodes.empty()) {
        fail(WOS_ERR_MESH, err);
        return -1;
    }
    for (size_t i = 0; i < b.nodes.size(); ++i) {
        if (count) count[i] = b.nodes[i].count;
        if (idx) idx[i] = b.nodes[i].idx;
        for (int k = 0; k < 3; ++k) {
            if (lo) lo[3 * i + k] = b.nodes[i].lo[k];
            if (hi) hi[3 * i + k] = b.nodes[i].hi[k];
        }
    }
    if (slots) std::memcpy(slots, b.slots.data(), b.slots.size() * sizeof(int32_t));
    return (int64_t)b.nodes.size();
}

int wos_ctx_set_tuning(wos_ctx *c, const wos_tuning *t) {
    if (!c || !t) return fail(WOS_ERR_INPUT, "context or tuning is NULL");
    if (t->block_threads < 64 || t->block_threads > kBlock || t->block_threads % 64)
        return fail(WOS_ERR_INPUT, "block_threads must be 64, 128, 192 or 256");
    if (t->chunk < 1 || t->chunk > (1 << 20)) return fail(WOS_ERR_INPUT, "chunk out of range");
    if (t->lds_nodes < -1 || t->blocks_per_cu < 0)
        return fail(WOS_ERR_INPUT, "lds_nodes must be >= -1 and blocks_per_cu >= 0");
    if (t->min_ready < 0 || t->min_ready > 64) return fail(WOS_ERR_INPUT, "min_ready must be in 0..64");
    if (!(t->relax >= 0.0f && t->relax < 0.5f)) return fail(WOS_ERR_INPUT, "relax must be in [0, 0.5)");
    if (t->grid_cells < -1 || t->grid_cells > (1 << 26))
        return fail(WOS_ERR_INPUT, "grid_cells must be -1 (default) or in 0..2^26");
    if (t->grid_list_max < 1 || t->grid_list_max > kGridListMax)
        return fail(WOS_ERR_INPUT, "grid_list_max must be in 1..127");
    if (t->xcd_parts < 1 || t->xcd_parts > kMaxParts)
        return fail(WOS_ERR_INPUT, "xcd_parts must be in 1..64");
    if (t->bvh_width != 2 && t->bvh_width != 4) return fail(WOS_ERR_INPUT, "bvh_width must be 2 or 4");
    if (t->stack_lds < 0 || t->stack_lds > 64) return fail(WOS_ERR_INPUT, "stack_lds must be in 0..64");
    c->tune = *t;
    return WOS_OK;
}

int wos_ctx_get_tuning(const wos_ctx *c, wos_tuning *t) {
    if (!c || !t) return fail(WOS_ERR_INPUT, "context or tuning is NULL");
    *t = c->tune;
    return WOS_OK;
}

int wos_ctx_autotune(wos_ctx *c, int nx, int ny, int nz, const wos_eq_desc *eq,
                     const wos_run_params *params, wos_tuning *chosen) {
    if (!c || !c->has_mesh) return fail(c ? WOS_ERR_STATE : WOS_ERR_INPUT, "no mesh set on this context");
    if (!params) return fail(WOS_ERR_INPUT, "params is NULL");
    if (nx < 1 || ny < 1 || nz < 1) return fail(WOS_ERR_INPUT, "grid sizes must be >= 1");
    if (c->mesh.dim == 2) nz = 1;
    GridSpec g{};
    g.nx = nx; g.ny = ny; g.nz = nz;
    std::vector<Part> parts;
    for (int d = 0; d < (int)c->L.domains.size(); ++d)
        parts.push_back(Part{d, &g, nullptr, (int64_t)nx * ny * nz});
    wos_tuning best{};
    const int rc = autotune(c, parts, eq, params, &best);
    if (rc != WOS_OK) return rc;
    c->tune = best;
    if (chosen) *chosen = best;
    return WOS_OK;
}

int wos_solve_grid(wos_ctx *c, int nx, int ny, int nz, const wos_eq_desc *eq,
                   const wos_run_params *params, wos_point_stats *out, wos_run_summary *summary) {
    if (!c || !c->has_mesh) return fail(c ? WOS_ERR_STATE : WOS_ERR_INPUT, "no mesh set on this context");
    if (nx < 1 || ny < 1 || nz < 1) return fail(WOS_ERR_INPUT, "grid sizes must be >= 1");
    if (c->mesh.dim == 2) nz = 1;  // runner.hpp:83
    GridSpec g{};
    g.nx = nx; g.ny = ny; g.nz = nz;
    return solve(c, {Part{0, &g, nullptr, (int64_t)nx * ny * nz}}, eq, params, out, summary);
}

int wos_solve_grid_batch(wos_ctx *c, int nx, int ny, int nz, const wos_eq_desc *eq,
                         const wos_run_params *params, wos_point_stats *out,
                         wos_run_summary *domain_summaries, wos_run_summary *summary) {
    if (!c || !c->has_mesh) return fail(c ? WOS_ERR_STATE : WOS_ERR_INPUT, "no mesh set on this context");
    if (nx < 1 || ny < 1 || nz < 1) return fail(WOS_ERR_INPUT, "grid sizes must be >= 1");
    if (c->mesh.dim == 2) nz = 1;
    GridSpec g{};
    g.nx = nx; g.ny = ny; g.nz = nz;
    std::vector<Part> parts;
    for (int d = 0; d < (int)c->L.domains.size(); ++d)
        parts.push_back(Part{d, &g, nullptr, (int64_t)nx * ny * nz});
    return solve(c, parts, eq, params, out, summary, nullptr, domain_summaries);
}

int wos_solve_points(wos_ctx *c, const double *points, int64_t n, const wos_eq_desc *eq,
                     const wos_run_params *params, wos_point_stats *out,
                     wos_run_summary *summary) {
    if (!points) return fail(WOS_ERR_INPUT, "points is NULL");
    return solve(c, {Part{0, nullptr, points, n}}, eq, params, out, summary);
}

int wos_solve_points_indexed(wos_ctx *c, const double *points, const int64_t *seed_index,
                             int64_t n, const wos_eq_desc *eq, const wos_run_params *params,
                             wos_point_stats *out, wos_run_summary *summary) {
    if (!points) return fail(WOS_ERR_INPUT, "points is NULL");
    if (!seed_index) return fail(WOS_ERR_INPUT, "seed_index is NULL");
    Part p{0, nullptr, points, n};
    p.seed_index = seed_index;
    return solve(c, {p}, eq, params, out, summary);
}

int wos_closest_points(wos_ctx *c, int domain, const double *points, int64_t n, int precision,
                       double relax, double *dist, double *lower, double *nearest,
                       int32_t *prim) {
    if (!c || !c->has_mesh) return fail(c ? WOS_ERR_STATE : WOS_ERR_INPUT, "no mesh set on this context");
    if (domain < 0 || domain >= (int)c->L.domains.size()) return fail(WOS_ERR_INPUT, "domain out of range");
    if (n < 0) return fail(WOS_ERR_INPUT, "n must be >= 0");
    if (n > 0 && (!points || !dist)) return fail(WOS_ERR_INPUT, "points and dist must not be NULL");
    if (precision != WOS_FP64_PLAIN && precision != WOS_FP32_PERSISTENT)
        return fail(WOS_ERR_INPUT, "unknown precision");
    if (!(relax >= 0.0 && relax < 0.5)) return fail(WOS_ERR_INPUT, "relax must be in [0, 0.5)");
    if (precision == WOS_FP64_PLAIN && relax != 0.0)
        return fail(WOS_ERR_INPUT, "the FP64 search is always exact (relax must be 0)");
    if (n == 0) return WOS_OK;
    const int dim = c->mesh.dim;
    for (int64_t i = 0; i < n * dim; ++i)
        if (!std::isfinite(points[i])) return fail(WOS_ERR_INPUT, "points must be finite");
    HIPX(hipSetDevice(c->device));
    std::vector<double> pts(points, points + n * dim);
    HIPX(upload(c->plist_in, pts, c->stream));
    HIPX(c->cq_out.reserve((size_t)n * (2 + dim) * sizeof(double)));
    HIPX(c->cq_prim.reserve((size_t)n * sizeof(int32_t)));
    double *d_dist = c->cq_out.as<double>(), *d_low = d_dist + n, *d_near = d_low + n;
    const DomainInfo &D = c->L.domains[domain];
    const float cf = (float)((1.0 - relax) * (1.0 - relax));
    const int blocks = (int)((n + kBlock - 1) / kBlock);
    const int width = c->tune.bvh_width;
    const size_t lds = (size_t)kBlock * std::max((c->L.depth + 2) * sizeof(int),
                                                 stack_entries(c, width) * sizeof(StackEntry));
    const int fp32 = precision == WOS_FP32_PERSISTENT;
    if (fp32) HIPX(ensure_grid(c));
    const int32_t root32 = width == 4 ? D.root_ref_w : D.root_ref;
    if (dim == 2 && width == 4)
        hipLaunchKernelGGL((k_closest<2, 4>), dim3(blocks), dim3(kBlock), lds, c->stream, c->view,
                           D.node_off, root32, domain, fp32, cf, c->plist_in.as<double>(), n, d_dist,
                           d_low, d_near, c->cq_prim.as<int32_t>());
    else if (dim == 2)
        hipLaunchKernelGGL((k_closest<2, 2>), dim3(blocks), dim3(kBlock), lds, c->stream, c->view,
                           D.node_off, root32, domain, fp32, cf, c->plist_in.as<double>(), n, d_dist,
                           d_low, d_near, c->cq_prim.as<int32_t>());
    else if (width == 4)
        hipLaunchKernelGGL((k_closest<3, 4>), dim3(blocks), dim3(kBlock), lds, c->stream, c->view,
                           D.node_off, root32, domain, fp32, cf, c->plist_in.as<double>(), n, d_dist,
                           d_low, d_near, c->cq_prim.as<int32_t>());
    else
        hipLaunchKernelGGL((k_closest<3, 2>), dim3(blocks), dim3(kBlock), lds, c->stream, c->view,
                           D.node_off, root32, domain, fp32, cf, c->plist_in.as<double>(), n, d_dist,
                           d_low, d_near, c->cq_prim.as<int32_t>());
    HIPX(hipGetLastError());
    HIPX(hipMemcpyAsync(dist, d_dist, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (lower)
        HIPX(hipMemcpyAsync(lower, d_low, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (nearest)
        HIPX(hipMemcpyAsync(nearest, d_near, n * dim * sizeof(double), hipMemcpyDeviceToHost,
                            c->stream));
    if (prim)
        HIPX(hipMemcpyAsync(prim, c->cq_prim.p, n * sizeof(int32_t), hipMemcpyDeviceToHost,
                            c->stream));
    HIPX(hipStreamSynchronize(c->stream));
    return WOS_OK;
}

int wos_closest_points_noalloc(wos_ctx *c, int domain, const double *points, int64_t n,
                               int precision, double relax, double *dist, double *lower,
                               double *nearest, int32_t *prim) {
    if (!c || !c->has_mesh) return fail(c ? WOS_ERR_STATE : WOS_ERR_INPUT, "no mesh set on this context");
    if (domain < 0 || domain >= (int)c->L.domains.size()) return fail(WOS_ERR_INPUT, "domain out of range");
    if (n < 0) return fail(WOS_ERR_INPUT, "n must be >= 0");
    if (n > 0 && (!points || !dist)) return fail(WOS_ERR_INPUT, "points and dist must not be NULL");
    if (precision != WOS_FP64_PLAIN && precision != WOS_FP32_PERSISTENT)
        return fail(WOS_ERR_INPUT, "unknown precision");
    if (!(relax >= 0.0 && relax < 0.5)) return fail(WOS_ERR_INPUT, "relax must be in [0, 0.5)");
    if (precision == WOS_FP64_PLAIN && relax != 0.0)
        return fail(WOS_ERR_INPUT, "the FP64 search is always exact (relax must be 0)");
    if (n == 0) return WOS_OK;
    const int dim = c->mesh.dim;
    // staging in the scratch array: points, then dist, lower, nearest, prim,
    // each part 256-byte aligned
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t b_pts = up((size_t)n * dim * sizeof(double)), b_d = up((size_t)n * sizeof(double));
    const size_t b_near = up((size_t)n * dim * sizeof(double)), b_prim = up((size_t)n * sizeof(int32_t));
    const int64_t per_point = (int64_t)(dim * 16 + 20);
    if ((double)n * (double)per_point + 5.0 * 256.0 > (double)kScratchBytes ||
        b_pts + 2 * b_d + b_near + b_prim > kScratchBytes)
        return fail(WOS_ERR_INPUT, "wos_closest_points_noalloc stages at most " +
                                       std::to_string((kScratchBytes - 5 * 256) / per_point) +
                                       " points (a fixed " + std::to_string(kScratchBytes >> 20) +
                                       " MiB device array); use wos_closest_points");
    for (int64_t i = 0; i < n * dim; ++i)
        if (!std::isfinite(points[i])) return fail(WOS_ERR_INPUT, "points must be finite");
    const int fp32 = precision == WOS_FP32_PERSISTENT;
    if (fp32) {
        if (c->tune.grid_cells == 0) {
            c->view.grids = nullptr; c->view.cells = nullptr; c->view.glist = nullptr;
        } else if (!c->grid_valid || c->grid_built_cells != c->tune.grid_cells ||
                   c->grid_built_cap != c->tune.grid_list_max ||
                   c->grid_built_relax != c->tune.relax) {
            return fail(WOS_ERR_STATE, "the FP32 distance grid of this mesh and tuning is not built "
                                       "(building it allocates): make one FP32 call on this "
                                       "context first, e.g. wos_closest_points");
        } else {
            c->view.grids = c->g_info.as<GridInfo>();
            c->view.cells = c->g_cells.as<GridCellF>();
            c->view.glist = c->g_list.as<GridEntry>();
        }
    }
    HIPX(hipSetDevice(c->device));
    unsigned char *base = static_cast<unsigned char *>(c->scratch_dev);
    double *d_pts = reinterpret_cast<double *>(base);
    double *d_dist = reinterpret_cast<double *>(base + b_pts);
    double *d_low = reinterpret_cast<double *>(base + b_pts + b_d);
    double *d_near = reinterpret_cast<double *>(base + b_pts + 2 * b_d);
    int32_t *d_prim = reinterpret_cast<int32_t *>(base + b_pts + 2 * b_d + b_near);
    HIPX(hipMemcpyAsync(d_pts, points, (size_t)n * dim * sizeof(double), hipMemcpyHostToDevice,
                        c->stream));
    const DomainInfo &D = c->L.domains[domain];
    const float cf = (float)((1.0 - relax) * (1.0 - relax));
    const int blocks = (int)((n + kBlock - 1) / kBlock);
    const int width = c->tune.bvh_width;
    const size_t lds = (size_t)kBlock * std::max((c->L.depth + 2) * sizeof(int),
                                                 stack_entries(c, width) * sizeof(StackEntry));
    const int32_t root32 = width == 4 ? D.root_ref_w : D.root_ref;
    if (dim == 2 && width == 4)
        hipLaunchKernelGGL((k_closest<2, 4>), dim3(blocks), dim3(kBlock), lds, c->stream, c->view,
                           D.node_off, root32, domain, fp32, cf, d_pts, n, d_dist, d_low, d_near,
                           d_prim);
    else if (dim == 2)
        hipLaunchKernelGGL((k_closest<2, 2>), dim3(blocks), dim3(kBlock), lds, c->stream, c->view,
                           D.node_off, root32, domain, fp32, cf, d_pts, n, d_dist, d_low, d_near,
                           d_prim);
    else if (width == 4)
        hipLaunchKernelGGL((k_closest<3, 4>), dim3(blocks), dim3(kBlock), lds, c->stream, c->view,
                           D.node_off, root32, domain, fp32, cf, d_pts, n, d_dist, d_low, d_near,
                           d_prim);
    else
        hipLaunchKernelGGL((k_closest<3, 2>), dim3(blocks), dim3(kBlock), lds, c->stream, c->view,
                           D.node_off, root32, domain, fp32, cf, d_pts, n, d_dist, d_low, d_near,
                           d_prim);
    HIPX(hipGetLastError());
    HIPX(hipMemcpyAsync(dist, d_dist, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (lower)
        HIPX(hipMemcpyAsync(lower, d_low, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (nearest)
        HIPX(hipMemcpyAsync(nearest, d_near, (size_t)n * dim * sizeof(double), hipMemcpyDeviceToHost,
                            c->stream));
    if (prim)
        HIPX(hipMemcpyAsync(prim, d_prim, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost,
                            c->stream));
    HIPX(hipStreamSynchronize(c->stream));
    return WOS_OK;
}

int wos_adjoint_points(wos_ctx *c, const double *points, int64_t n, const wos_eq_desc *eq,
                       const wos_run_params *params, const double *dl_du, double *grad_vertex,
                       double *grad_params, wos_point_stats *out, wos_run_summary *summary) {
    if (!points) return fail(WOS_ERR_INPUT, "points is NULL");
    AdjRequest r{dl_du, grad_vertex, grad_params};
    return solve(c, {Part{0, nullptr, points, n}}, eq, params, out, summary, &r);
}

int wos_adjoint_grid(wos_ctx *c, int nx, int ny, int nz, const wos_eq_desc *eq,
                     const wos_run_params *params, const double *dl_du, double *grad_vertex,
                     double *grad_params, wos_point_stats *out, wos_run_summary *summary) {
    if (!c || !c->has_mesh) return fail(c ? WOS_ERR_STATE : WOS_ERR_INPUT, "no mesh set on this context");
    if (nx < 1 || ny < 1 || nz < 1) return fail(WOS_ERR_INPUT, "grid sizes must be >= 1");
    if (c->mesh.dim == 2) nz = 1;
    GridSpec g{};
    g.nx = nx; g.ny = ny; g.nz = nz;
    AdjRequest r{dl_du, grad_vertex, grad_params};
    return solve(c, {Part{0, &g, nullptr, (int64_t)nx * ny * nz}}, eq, params, out, summary, &r);
}

}  // extern "C"
