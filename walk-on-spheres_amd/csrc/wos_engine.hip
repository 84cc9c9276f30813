// wos_engine.hip -- HIP kernels (gfx950) and the C ABI of include/wos_hip.h.
//
// Pipeline of one solve (replacing the body of run<N,Eq>,
// include/wos/runner.hpp:82-238):
//   1. k_classify: per point, FP64 start query (reference traversal), the
//      reference boundary / winding / ray-vote classification, the walk-stream
//      key and g at the nearest point;
//   2. host compaction of the inside points;
//   3. k_walk_fp64 (one lane per walk) or k_walk_fp32 (persistent lanes,
//      ballot-driven refill) writing one FP64 value per walk into a
//      [walk][point] staging buffer;
//   4. k_welford: per point, the reference's sequential Welford update in walk
//      order (include/wos/wos.hpp:169-180), continued across rounds.
// Results are bitwise reproducible: a walk's draws depend only on (point seed,
// walk, step), and the merge order is fixed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/wos_hip.h"
#include "bvh_build.hpp"
#include "wos_common.hpp"
#include "wos_device.hpp"

using namespace wosx;

namespace {

constexpr int kBlock = 256;

// --------------------------------------------------------------------------
// walk-step helpers (FP64): src/wos.cpp:9-56, include/wos/wos.hpp:54-138
// --------------------------------------------------------------------------
struct CursorD {
    Cursor c;
    __device__ CursorD(uint32_t w, uint32_t s, uint32_t tag, uint32_t k0, uint32_t k1)
        : c(w, s, tag, k0, k1) {}
    __device__ double unit() { return c.unit(); }
    __device__ double unit_open() { return c.unit_open(); }
};
struct CursorF {
    Cursor c;
    __device__ CursorF(uint32_t w, uint32_t s, uint32_t tag, uint32_t k0, uint32_t k1)
        : c(w, s, tag, k0, k1) {}
    __device__ float unit() { return c.unitf(); }
    __device__ float unit_open() { return c.unitf_open(); }
};

template <typename T> __device__ __forceinline__ T tsin(T x) { return sin(x); }
template <typename T> __device__ __forceinline__ T tcos(T x) { return cos(x); }

// include/wos/poisson.hpp:27-30, :46-50
template <typename T, typename RNG>
__device__ __forceinline__ T poisson_radius_2d(T R, RNG &rng) {
    const T u = rng.unit_open();
    const T v = rng.unit_open();
    return R * sqrt(u * v);
}
template <typename T, typename RNG>
__device__ __forceinline__ T poisson_radius_3d(T R, RNG &rng) {
    const T u1 = rng.unit_open(), u2 = rng.unit_open();
    const T a = -log(u1 * u2);
    const T v1 = rng.unit_open(), v2 = rng.unit_open();
    const T b = -log(v1 * v2);
    return R * a / (a + b);
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
// include/wos/screened.hpp:105-129
template <typename T, typename RNG>
__device__ T screened_radius_2d(T alpha, T R, RNG &rng) {
    const T X = alpha * R;
    for (int guard = 0; guard < 100000; ++guard) {
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
            const T log_ratio = log_i0<T>(x) + log_k0<T>(X) - log_i0<T>(X) - log_k0<T>(x);
            acceptance = -expm1(log_ratio < (T)0 ? log_ratio : (T)0);
        }
        if (rng.unit() < clamp01<T>(acceptance)) return r;
    }
    return (T)0.5 * R;
}
// include/wos/screened.hpp:131-156
template <typename T, typename RNG>
__device__ T screened_radius_3d(T alpha, T R, RNG &rng) {
    const T X = alpha * R;
    for (int guard = 0; guard < 100000; ++guard) {
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
    return (T)0.5 * R;
}

// One source term of wos_sample (include/wos/wos.hpp:69-110): returns the
// value to add (already multiplied by the weight).  T is the geometry type.
template <int DIM, typename T, typename RNG>
__device__ __forceinline__ double source_term(const EqView &e, T px, T py, T pz, T R, T weight,
                                              RNG &rng) {
    const bool screened = e.kind == WOS_EQ_SCREENED;
    const T alpha = (T)e.alpha;
    if (DIM == 2 && !eq_source_may_intersect<T>(e, px, py, R)) return 0.0;
    T sx, sy, sz = (T)0;
    if (e.source_mode == WOS_SOURCE_GREEN) {
        if (DIM == 2) {
            const T radius = screened ? screened_radius_2d<T>(alpha, R, rng) : poisson_radius_2d<T>(R, rng);
            const T angle = (T)2 * (T)kPi * rng.unit();
            sx = px + radius * cos(angle);
            sy = py + radius * sin(angle);
        } else {
            const T radius = screened ? screened_radius_3d<T>(alpha, R, rng) : poisson_radius_3d<T>(R, rng);
            const T z = (T)2 * rng.unit() - (T)1;
            const T angle = (T)2 * (T)kPi * rng.unit();
            const T xy = sqrt((T)1 - z * z);
            sx = px + radius * xy * cos(angle);
            sy = py + radius * xy * sin(angle);
            sz = pz + radius * z;
        }
        const T f = eq_source<T>(e, sx, sy, sz, DIM);
        if (f == (T)0) return 0.0;
        T mass;
        if (DIM == 2) mass = screened ? screened_mass_2d<T>(alpha, R) : R * R / (T)4;
        else mass = screened ? screened_mass_3d<T>(alpha, R) : R * R / (T)6;
        return (double)(weight * mass * f);
    }
    // uniform ball point (src/wos.cpp:33-56)
    T dist;
    if (DIM == 2) {
        const T theta = rng.unit() * (T)2 * (T)kPi;
        const T r = R * sqrt(rng.unit_open());
        sx = px + r * cos(theta);
        sy = py + r * sin(theta);
        const T dx = px - sx, dy = py - sy;
        dist = sqrt(dx * dx + dy * dy);
    } else {
        const T u = rng.unit(), v = rng.unit(), w = rng.unit_open();
        const T z = (T)2 * u - (T)1, phi = (T)2 * (T)kPi * v;
        const T surf = sqrt((T)1 - z * z), vol = R * cbrt(w);
        sx = px + vol * surf * cos(phi);
        sy = py + vol * surf * sin(phi);
        sz = pz + vol * z;
        const T dx = px - sx, dy = py - sy, dz = pz - sz;
        dist = sqrt(dx * dx + dy * dy + dz * dz);
    }
    const T f = eq_source<T>(e, sx, sy, sz, DIM);
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

__device__ __forceinline__ double grid_coordinate(double lo, double hi, int index, int count) {
    return count > 1 ? lo + (hi - lo) * (double)index / (double)(count - 1) : lo;
}

template <int DIM>
__global__ __launch_bounds__(kBlock) void k_classify(MeshView m, EqView e, GridSpec g,
                                                     const double *plist, int64_t n, uint64_t seed,
                                                     double tol, int max_ray_attempts,
                                                     uint8_t *loc, StartInfo *starts,
                                                     int32_t *amb, double *bvalue) {
    extern __shared__ int lds_i[];
    int *stk = lds_i + threadIdx.x;
    const int stride = blockDim.x;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x, y, z = 0.0;
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
    const uint64_t ps = point_seed(seed, (uint64_t)i);
    NpqD q;
    if (DIM == 2) q = npq_ref_2d(m, x, y, stk, stride);
    else q = npq_ref_3d(m, x, y, z, stk, stride);
    const double r = sqrt(q.d2);
    int code;
    int retries = 0;
    if (r <= tol) {
        code = 2;
    } else if (DIM == 2) {
        code = winding_number_d(m, x, y) != 0 ? 1 : 0;
    } else {
        // src/inside.cpp:228-260 with the inside stream of runner.hpp:182-183
        Xoshiro rng(splitmix64(ps ^ kInsideStream));
        int valid = 0, votes = 0;
        for (int attempts = 0; attempts < max_ray_attempts && valid < 3; ++attempts) {
            const double u = rng.unit(), v = rng.unit();
            const double zz = 2 * u - 1;
            const double phi = 2 * kPi * v;
            const double rr = sqrt(1 - zz * zz);
            bool ambiguous;
            const int hits = ray_cast_d(m, x, y, z, rr * cos(phi), rr * sin(phi), zz, ambiguous,
                                        stk, stride);
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
    S.pad = 0;
    S.bid = (DIM == 2 && q.prim >= 0) ? m.prim_bid[q.prim] : -1;
    S.key0 = (uint32_t)key;
    S.key1 = (uint32_t)(key >> 32);
    starts[i] = S;
    loc[i] = (uint8_t)code;
    amb[i] = retries;
    bvalue[i] = gval;
}

// --------------------------------------------------------------------------
// one round of walks: active point k runs walks [wbeg[k], wbeg[k] + count_k);
// item i of the round is walk (i - off[k]) of the point with
// off[k] <= i < off[k+1], and its result goes to staging slot i (point major).
// off == nullptr: every point runs cw walks starting at w0.
// --------------------------------------------------------------------------
struct RoundView {
    const StartInfo *starts;
    const int32_t *inside;   // inside id -> point index
    const int32_t *pid;      // active k -> inside id
    const int32_t *wbeg;     // active k -> first walk index (off != nullptr)
    const int64_t *off;      // exclusive prefix of counts, n + 1 entries, or nullptr
    int n;
    int cw;
    int w0;
    int64_t total;
};

__device__ __forceinline__ const StartInfo &round_item(const RoundView &r, int64_t i, uint32_t &walk) {
    int k;
    if (!r.off) {
        k = (int)(i / r.cw);
        walk = (uint32_t)(r.w0 + (int)(i - (int64_t)k * r.cw));
    } else {
        int lo = 0, hi = r.n - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (r.off[mid] <= i) lo = mid; else hi = mid - 1;
        }
        k = lo;
        walk = (uint32_t)(r.wbeg[k] + (int)(i - r.off[k]));
    }
    return r.starts[r.inside[r.pid[k]]];
}

// --------------------------------------------------------------------------
// kernel 2a: FP64 plain walker, one lane per (point, walk)
// --------------------------------------------------------------------------
template <int DIM, bool SRC, bool SCR>
__global__ __launch_bounds__(kBlock) void k_walk_fp64(MeshView m, EqView e, RoundView rv,
                                                      double eps, int max_steps, double *vals,
                                                      int32_t *steps_out) {
    extern __shared__ int lds_i[];
    int *stk = lds_i + threadIdx.x;
    const int stride = blockDim.x;
    const int64_t item = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= rv.total) return;
    uint32_t w;
    const StartInfo S = round_item(rv, item, w);
    double px = S.x, py = S.y, pz = S.z, R = S.r;
    NpqD q;
    q.px = S.nx; q.py = S.ny; q.pz = S.nz; q.s = S.s; q.t = S.t; q.prim = S.prim;
    double sample = 0.0, weight = 1.0;
    int steps = 0;
    const double alpha = e.alpha;
    while (R > eps && steps < max_steps) {
        if (SRC && e.source_kind != WOS_SRC_ZERO) {
            CursorD src(w, (uint32_t)steps, 1, S.key0, S.key1);
            sample += source_term<DIM, double>(e, px, py, pz, R, weight, src);
        }
        if (SCR) weight *= DIM == 2 ? screened_weight_2d<double>(alpha, R)
                                    : screened_weight_3d<double>(alpha, R);
        CursorD wk(w, (uint32_t)steps, 0, S.key0, S.key1);
        if (DIM == 2) {
            const double theta = wk.unit() * 2 * kPi;
            px = px + R * cos(theta);
            py = py + R * sin(theta);
        } else {
            const double u = wk.unit(), v = wk.unit();
            const double z = 2 * u - 1, phi = 2 * kPi * v;
            const double rr = sqrt(1 - z * z);
            px = px + R * rr * cos(phi);
            py = py + R * rr * sin(phi);
            pz = pz + R * z;
        }
        ++steps;
        q = DIM == 2 ? npq_ref_2d(m, px, py, stk, stride) : npq_ref_3d(m, px, py, pz, stk, stride);
        R = sqrt(q.d2);
    }
    const bool hit = R > eps;
    const double g = steps == 0 ? S.g : eq_boundary(e, m, q.px, q.py, q.pz, q.prim, q.s, q.t);
    vals[item] = sample + weight * g;
    steps_out[item] = steps | (hit ? (int32_t)0x80000000 : 0);
}

// FP32 source term.  Same draws as source_term (the walk's tag-1 counter
// block), but the Green function takes the sample's distance from the
// centre in closed form (R sqrt(u) in 2D, R cbrt(w) in 3D) rather than from
// rounded coordinates, and a constant source skips the direction entirely:
// only the radius draw is made (block 1 word 0 in 3D, block 0 word 1 in 2D).
template <int DIM, int SMODE>
__device__ __forceinline__ double source_term_f(const EqView &e, float px, float py, float pz,
                                                float R, float weight, uint32_t w, uint32_t step,
                                                uint32_t k0, uint32_t k1) {
    if (SMODE == 2) {
        CursorF src(w, step, 1, k0, k1);
        return source_term<DIM, float>(e, px, py, pz, R, weight, src);
    }
    if (DIM == 2 && !eq_source_may_intersect<float>(e, px, py, R)) return 0.0;
    const bool screened = e.kind == WOS_EQ_SCREENED;
    const float alpha = (float)e.alpha;
    float f, dist;
    if (e.source_kind == WOS_SRC_CONSTANT) {
        if (DIM == 2) {
            const U64x2 r = philox(w, step, 0, 1, k0, k1);
            dist = R * sqrtf(u01f_open(r.b));
        } else {
            const U64x2 r = philox(w, step, 1, 1, k0, k1);
            dist = R * cbrtf(u01f_open(r.a));
        }
        f = (float)e.src_value;
    } else {
        CursorF rng(w, step, 1, k0, k1);
        float sx, sy, sz = 0.0f;
        if (DIM == 2) {
            const float theta = rng.unit();
            dist = R * sqrtf(rng.unit_open());
            float sn, cs;
            __sincosf(theta * (float)(2.0 * kPi), &sn, &cs);
            sx = __builtin_fmaf(dist, cs, px);
            sy = __builtin_fmaf(dist, sn, py);
        } else {
            const float u = rng.unit(), v = rng.unit(), ww = rng.unit_open();
            const float z = 2.0f * u - 1.0f;
            const float surf = sqrtf(fmaxf(0.0f, 1.0f - z * z));
            dist = R * cbrtf(ww);
            float sn, cs;
            __sincosf(v * (float)(2.0 * kPi), &sn, &cs);
            sx = __builtin_fmaf(dist * surf, cs, px);
            sy = __builtin_fmaf(dist * surf, sn, py);
            sz = __builtin_fmaf(dist, z, pz);
        }
        f = eq_source<float>(e, sx, sy, sz, DIM);
    }
    if (f == 0.0f || !(dist > 0.0f)) return 0.0;
    float volume, G;
    if (DIM == 2) {
        volume = (float)kPi * R * R;
        G = screened ? screened_green_2d<float>(alpha, R, dist) : poisson_green_2d<float>(R, dist);
    } else {
        volume = (float)(4.0 * kPi / 3.0) * R * R * R;
        G = screened ? screened_green_3d<float>(alpha, R, dist) : poisson_green_3d<float>(R, dist);
    }
    return (double)(weight * volume * f * G);
}

// --------------------------------------------------------------------------
// kernel 2b: FP32 persistent walker with ballot-driven lane refill
// --------------------------------------------------------------------------
struct PersistArgs {
    RoundView rv;
    int max_steps;
    double eps;
    float eps_f;
    float shrink_rel;
    uint64_t total;
    unsigned long long *counter;
    int chunk;
    int n_cached;        // internal BVH nodes staged in LDS (breadth-first prefix)
    int cache_bytes;     // LDS bytes reserved for them (16 B aligned)
    double *vals;
    int32_t *steps_out;
};

template <int DIM, int SMODE, bool SCR>
__global__ __launch_bounds__(kBlock) void k_walk_fp32(MeshView m, EqView e, PersistArgs a) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    // stage the top of the tree once per block
    {
        const int4 *src = DIM == 3 ? reinterpret_cast<const int4 *>(m.nf3)
                                   : reinterpret_cast<const int4 *>(m.nf2);
        int4 *dst = reinterpret_cast<int4 *>(lds_raw);
        const int n16 = a.n_cached * (int)(DIM == 3 ? sizeof(NodeF3) : sizeof(NodeF2)) / 16;
        for (int i = threadIdx.x; i < n16; i += blockDim.x) dst[i] = src[i];
        __syncthreads();
    }
    const void *cache = lds_raw;
    StackEntry *stk = reinterpret_cast<StackEntry *>(lds_raw + a.cache_bytes) + threadIdx.x;
    const int stride = blockDim.x;
    const int lane = __lane_id();

    // lane state
    bool active = false;
    float px = 0, py = 0, pz = 0, R = 0, weight = 1.0f;
    double sample = 0.0;
    int steps = 0, slot = -1;
    uint32_t w = 0, k0 = 0, k1 = 0;
    size_t out_index = 0;
    const float alpha = (float)e.alpha;

    // wave-uniform pool of work items
    uint64_t pool_next = 0, pool_end = 0;
    bool exhausted = false;

    for (;;) {
        // ---- refill free lanes from the wave pool
        const uint64_t free_mask = __ballot(!active);
        if (free_mask != 0 && !exhausted) {
            if (pool_next >= pool_end) {
                unsigned long long base = 0;
                if (lane == 0) base = atomicAdd(a.counter, (unsigned long long)a.chunk);
                base = __shfl(base, 0);
                if (base >= (unsigned long long)a.rv.total) {
                    exhausted = true;
                } else {
                    pool_next = base;
                    pool_end = base + (uint64_t)a.chunk;
                    if (pool_end > (uint64_t)a.rv.total) pool_end = (uint64_t)a.rv.total;
                }
            }
            if (!exhausted) {
                const uint32_t lo = (uint32_t)free_mask, hi = (uint32_t)(free_mask >> 32);
                const int rank = __builtin_amdgcn_mbcnt_hi(hi, __builtin_amdgcn_mbcnt_lo(lo, 0));
                const uint64_t item = pool_next + (uint64_t)rank;
                if (!active && item < pool_end) {
                    const StartInfo &S = round_item(a.rv, (int64_t)item, w);
                    px = (float)S.x; py = (float)S.y; pz = (float)S.z;
                    R = (float)S.r * a.shrink_rel - m.shrink_abs;
                    weight = 1.0f;
                    sample = 0.0;
                    steps = 0;
                    slot = S.slot;
                    k0 = S.key0;
                    k1 = S.key1;
                    out_index = (size_t)item;
                    active = true;
                    if (!(R > a.eps_f)) {
                        // zero-step walk: g at the start's nearest point
                        a.vals[out_index] = S.g;
                        a.steps_out[out_index] = 0;
                        active = false;
                    }
                }
                const uint64_t nfree = (uint64_t)__popcll(free_mask);
                pool_next = pool_next + nfree < pool_end ? pool_next + nfree : pool_end;
            }
        }
        if (__ballot(active) == 0) {
            if (exhausted) break;
            continue;
        }
        if (active) {
        // ---- one walk step (include/wos/wos.hpp:68-127)
        if (SMODE != 0 && e.source_kind != WOS_SRC_ZERO)
            sample += source_term_f<DIM, SMODE>(e, px, py, pz, R, weight, w, (uint32_t)steps, k0, k1);
        if (SCR) weight *= DIM == 2 ? screened_weight_2d<float>(alpha, R)
                                    : screened_weight_3d<float>(alpha, R);
        Cursor wk(w, (uint32_t)steps, 0, k0, k1);
        if (DIM == 2) {
            const float u = wk.unitf();
            float sn, cs;
            __sincosf(u * (float)(2.0 * kPi), &sn, &cs);
            px = __builtin_fmaf(R, cs, px);
            py = __builtin_fmaf(R, sn, py);
        } else {
            const float u = wk.unitf(), v = wk.unitf();
            const float z = 2.0f * u - 1.0f;
            const float rr = sqrtf(fmaxf(0.0f, 1.0f - z * z));
            float sn, cs;
            __sincosf(v * (float)(2.0 * kPi), &sn, &cs);
            px = __builtin_fmaf(R * rr, cs, px);
            py = __builtin_fmaf(R * rr, sn, py);
            pz = __builtin_fmaf(R, z, pz);
        }
        ++steps;
        const float d2 = npq_f<DIM>(m, px, py, pz, slot, stk, stride, cache, a.n_cached);
        R = sqrtf(d2) * a.shrink_rel - m.shrink_abs;
        if (!(R > a.eps_f) || steps >= a.max_steps) {
            // terminal value in FP64: nearest point on the nearest primitive
            const bool hit = R > a.eps_f;
            double g = 0.0;
            if (slot >= 0) {
                const int prim = m.slot_prim[slot];
                double nx, ny, nz = 0.0, s, t = 0.0;
                if (DIM == 2) npq_seg_d(m.segd[slot], (double)px, (double)py, nx, ny, s);
                else eberly_d(m.trid[slot], (double)px, (double)py, (double)pz, nx, ny, nz, s, t);
                g = eq_boundary(e, m, nx, ny, nz, prim, s, t);
            }
            a.vals[out_index] = sample + (double)weight * g;
            a.steps_out[out_index] = steps | (hit ? (int32_t)0x80000000 : 0);
            active = false;
        }
        }  // active
    }
}

// --------------------------------------------------------------------------
// kernel 3: sequential Welford per point (include/wos/wos.hpp:169-180)
// --------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_welford(RoundView rv, const double *vals,
                                                    const int32_t *steps, double *mean, double *m2,
                                                    double *count, double *steps_sum,
                                                    int32_t *hits) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= rv.n) return;
    const int p = rv.pid[k];
    const int64_t base = rv.off ? rv.off[k] : (int64_t)k * rv.cw;
    const int cnt = rv.off ? (int)(rv.off[k + 1] - rv.off[k]) : rv.cw;
    double mu = mean[p], M2 = m2[p], n = count[p], ts = steps_sum[p];
    int h = hits[p];
    for (int j = 0; j < cnt; ++j) {
        const size_t o = (size_t)(base + j);
        const double x = vals[o];
        const int32_t s = steps[o];
        n = n + 1.0;
        const double delta_before = x - mu;
        mu += delta_before / n;
        const double delta_after = x - mu;
        M2 += delta_before * delta_after;
        ts += (double)(s & 0x7fffffff);
        h += (s < 0) ? 1 : 0;
    }
    mean[p] = mu; m2[p] = M2; count[p] = n; steps_sum[p] = ts; hits[p] = h;
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

template <typename T>
hipError_t upload(DevBuf &b, const std::vector<T> &v, hipStream_t s) {
    hipError_t e = b.reserve(v.size() * sizeof(T));
    if (e != hipSuccess || v.empty()) return e;
    e = hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s);
}

}  // namespace

struct wos_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int n_cu = 0;
    HostMesh mesh;
    HostBVH bvh;
    DeviceLayouts L;
    bool has_mesh = false;
    DevBuf nd2, nd3, segd, trid, slot_prim, prim_bid, prims, verts, nf2, nf3, segf, seg_inv, trif;
    DevBuf plist_in, loc, starts, amb, bvalue, inside, vals, steps, mean, m2, count, ssum, hits,
        counter, vv, pid_all, r_pid, r_wbeg, r_off;
    int last_rounds = 0;
    MeshView view{};
    hipEvent_t ev[6]{};
    wos_tuning tune{256, -1, 256, 0};
    size_t lds_per_cu = 65536;
    size_t lds_per_block = 65536;
    int last_threads = 0;
    int last_cached = 0;
};

namespace {

template <typename K>
hipError_t launch_classify(wos_ctx *c, K kernel, const EqView &ev, const GridSpec &g,
                           const double *plist, int64_t n, uint64_t seed, double tol, int mra) {
    const int blocks = (int)((n + kBlock - 1) / kBlock);
    const size_t lds = (size_t)kBlock * (c->bvh.depth + 2) * sizeof(int);
    hipLaunchKernelGGL(kernel, dim3(blocks), dim3(kBlock), lds, c->stream, c->view, ev, g, plist,
                       n, seed, tol, mra, c->loc.as<uint8_t>(), c->starts.as<StartInfo>(),
                       c->amb.as<int32_t>(), c->bvalue.as<double>());
    return hipGetLastError();
}

template <int DIM, bool SRC, bool SCR>
hipError_t launch_walk(wos_ctx *c, const EqView &ev, int precision, const RoundView &rv,
                       double eps, int max_steps, int *blocks_out) {
    if (precision == WOS_FP64_PLAIN) {
        const int blocks = (int)((rv.total + kBlock - 1) / kBlock);
        const size_t lds = (size_t)kBlock * (c->bvh.depth + 2) * sizeof(int);
        hipLaunchKernelGGL((k_walk_fp64<DIM, SRC, SCR>), dim3(blocks), dim3(kBlock), lds,
                           c->stream, c->view, ev, rv, eps, max_steps, c->vals.as<double>(),
                           c->steps.as<int32_t>());
        *blocks_out = blocks;
        return hipGetLastError();
    }
    constexpr int SMODE = SRC ? 1 : 0;
    const void *kfn = ev.source_mode == WOS_SOURCE_GREEN && SRC
                          ? (const void *)k_walk_fp32<DIM, 2, SCR>
                          : (const void *)k_walk_fp32<DIM, SMODE, SCR>;
    const int threads = c->tune.block_threads;
    const size_t stack_bytes = (size_t)threads * (c->bvh.depth + 1) * sizeof(StackEntry);
    const int node_bytes = DIM == 3 ? (int)sizeof(NodeF3) : (int)sizeof(NodeF2);
    const int n_internal = DIM == 3 ? (int)c->L.nf3.size() : (int)c->L.nf2.size();
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
    const size_t lds = stack_bytes + cache_bytes;
    if (lds > c->lds_per_block) return hipErrorInvalidValue;
    int per_cu2 = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu2, kfn, threads, lds);
    if (e != hipSuccess) return e;
    per_cu2 = std::max(1, std::min(per_cu2, per_cu));
    const int blocks = c->n_cu * per_cu2;
    PersistArgs a;
    a.n_cached = n_cached;
    a.cache_bytes = cache_bytes;
    a.rv = rv;
    a.max_steps = max_steps;
    a.eps = eps;
    a.eps_f = (float)eps;
    a.shrink_rel = 1.0f - 1.0f / 131072.0f;
    a.counter = c->counter.as<unsigned long long>();
    a.chunk = c->tune.chunk;
    a.vals = c->vals.as<double>();
    a.steps_out = c->steps.as<int32_t>();
    e = hipMemsetAsync(c->counter.p, 0, sizeof(unsigned long long), c->stream);
    if (e != hipSuccess) return e;
    if (ev.source_mode == WOS_SOURCE_GREEN && SRC)
        hipLaunchKernelGGL((k_walk_fp32<DIM, 2, SCR>), dim3(blocks), dim3(threads), lds,
                           c->stream, c->view, ev, a);
    else
        hipLaunchKernelGGL((k_walk_fp32<DIM, SMODE, SCR>), dim3(blocks), dim3(threads), lds,
                           c->stream, c->view, ev, a);
    *blocks_out = blocks;
    c->last_threads = threads;
    c->last_cached = n_cached;
    return hipGetLastError();
}

template <int DIM>
hipError_t dispatch_walk(wos_ctx *c, const EqView &ev, bool src, bool scr, int precision,
                         const RoundView &rv, double eps, int max_steps, int *blocks) {
    if (src && scr) return launch_walk<DIM, true, true>(c, ev, precision, rv, eps, max_steps, blocks);
    if (src) return launch_walk<DIM, true, false>(c, ev, precision, rv, eps, max_steps, blocks);
    if (scr) return launch_walk<DIM, false, true>(c, ev, precision, rv, eps, max_steps, blocks);
    return launch_walk<DIM, false, false>(c, ev, precision, rv, eps, max_steps, blocks);
}

int validate_eq(const wos_ctx *c, const wos_eq_desc *eq) {
    if (!eq) return fail(WOS_ERR_INPUT, "equation descriptor is NULL");
    if (eq->kind < 0 || eq->kind > 2) return fail(WOS_ERR_INPUT, "unknown equation kind");
    if (eq->source_mode < 0 || eq->source_mode > 1) return fail(WOS_ERR_INPUT, "unknown source mode");
    if (eq->source_kind < 0 || eq->source_kind > 5) return fail(WOS_ERR_INPUT, "unknown source kind");
    if (eq->boundary_kind < 0 || eq->boundary_kind > 5) return fail(WOS_ERR_INPUT, "unknown boundary kind");
    if (eq->kind == WOS_EQ_LAPLACE && eq->source_kind != WOS_SRC_NONE)
        return fail(WOS_ERR_INPUT, "laplace takes no source");
    if (eq->kind == WOS_EQ_SCREENED && !(eq->alpha > 0.0 && std::isfinite(eq->alpha)))
        return fail(WOS_ERR_INPUT, "screened_poisson needs a positive finite alpha");
    if (eq->source_kind == WOS_SRC_GAUSSIAN && !(eq->src_radius > 0.0))
        return fail(WOS_ERR_INPUT, "gaussian source needs a positive width");
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

// shared body of wos_solve_grid / wos_solve_points
int solve(wos_ctx *c, const GridSpec *grid, const double *points, int64_t n,
          const wos_eq_desc *eq, const wos_run_params *prm, wos_point_stats *out,
          wos_run_summary *summary) {
    if (!c) return fail(WOS_ERR_INPUT, "context is NULL");
    if (!c->has_mesh) return fail(WOS_ERR_STATE, "no mesh set on this context");
    if (!prm || !out || !out->mean) return fail(WOS_ERR_INPUT, "params and out->mean are required");
    int rc = validate_eq(c, eq);
    if (rc) return rc;
    if (prm->n_walks < 1) return fail(WOS_ERR_INPUT, "n_walks must be >= 1");
    if (prm->max_steps < 1) return fail(WOS_ERR_INPUT, "max_steps must be >= 1");
    if (c->mesh.dim == 3 && prm->max_ray_attempts < 3)
        return fail(WOS_ERR_INPUT, "max_ray_attempts must be >= 3");
    if (!(prm->epsilon > 0.0) || !std::isfinite(prm->epsilon))
        return fail(WOS_ERR_INPUT, "epsilon must be positive and finite");
    if (prm->precision != WOS_FP64_PLAIN && prm->precision != WOS_FP32_PERSISTENT)
        return fail(WOS_ERR_INPUT, "unknown precision");
    if (n <= 0 || n > ((int64_t)1 << 31) - 1) return fail(WOS_ERR_INPUT, "point count out of range");
    if (!(prm->target_se >= 0.0) || !std::isfinite(prm->target_se))
        return fail(WOS_ERR_INPUT, "target_se must be finite and >= 0");
    if (prm->target_se > 0.0 && prm->max_walks < prm->n_walks)
        return fail(WOS_ERR_INPUT, "adaptive mode needs max_walks >= n_walks");
    const int dim = c->mesh.dim;
    HIPX(hipSetDevice(c->device));

    EqView ev{};
    ev.kind = eq->kind; ev.source_mode = eq->source_mode; ev.source_kind = eq->source_kind;
    ev.boundary_kind = eq->boundary_kind; ev.alpha = eq->alpha; ev.src_value = eq->src_value;
    ev.cx = eq->src_center[0]; ev.cy = eq->src_center[1]; ev.cz = eq->src_center[2];
    ev.src_radius = eq->src_radius;
    for (int i = 0; i < WOS_MAX_BND_VALUES; ++i) ev.bnd[i] = eq->bnd_values[i];
    ev.n_bnd = eq->n_bnd_values;
    ev.vertex_values = nullptr;
    if (eq->boundary_kind == WOS_BND_VERTEX) {
        std::vector<double> vv(eq->vertex_values, eq->vertex_values + eq->n_vertex_values);
        HIPX(upload(c->vv, vv, c->stream));
        ev.vertex_values = c->vv.as<double>();
    }

    // bounding box and tolerance (runner.hpp:82-88)
    double gb[6];
    {
        const auto &V = c->mesh.verts;
        double lo[3] = {V[0], V[1], dim == 3 ? V[2] : 0.0};
        double hi[3] = {lo[0], lo[1], lo[2]};
        for (int64_t i = 1; i < c->mesh.n_verts(); ++i)
            for (int k = 0; k < dim; ++k) {
                const double v = V[(size_t)i * dim + k];
                if (v < lo[k]) lo[k] = v;
                else if (v > hi[k]) hi[k] = v;
            }
        gb[0] = lo[0]; gb[1] = hi[0]; gb[2] = lo[1]; gb[3] = hi[1]; gb[4] = lo[2]; gb[5] = hi[2];
    }
    const double scale = std::max({1.0, gb[1] - gb[0], gb[3] - gb[2], gb[5] - gb[4]});
    const double tol = 1e-10 * scale;

    GridSpec g{};
    if (grid) {
        g = *grid;
        for (int k = 0; k < 6; ++k) g.b[k] = gb[k];
    }

    HIPX(c->loc.reserve(n));
    HIPX(c->starts.reserve(n * sizeof(StartInfo)));
    HIPX(c->amb.reserve(n * sizeof(int32_t)));
    HIPX(c->bvalue.reserve(n * sizeof(double)));
    const double *dpts = nullptr;
    if (points) {
        HIPX(c->plist_in.reserve(n * dim * sizeof(double)));
        HIPX(hipMemcpyAsync(c->plist_in.p, points, n * dim * sizeof(double), hipMemcpyHostToDevice,
                            c->stream));
        dpts = c->plist_in.as<double>();
    }

    HIPX(hipEventRecord(c->ev[0], c->stream));
    hipError_t le = dim == 2 ? launch_classify(c, k_classify<2>, ev, g, dpts, n, prm->seed, tol,
                                               prm->max_ray_attempts)
                             : launch_classify(c, k_classify<3>, ev, g, dpts, n, prm->seed, tol,
                                               prm->max_ray_attempts);
    HIPX(le);
    HIPX(hipEventRecord(c->ev[1], c->stream));

    std::vector<uint8_t> loc(n);
    std::vector<int32_t> amb(n);
    std::vector<double> bval(n);
    HIPX(hipMemcpyAsync(loc.data(), c->loc.p, n, hipMemcpyDeviceToHost, c->stream));
    HIPX(hipMemcpyAsync(amb.data(), c->amb.p, n * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIPX(hipMemcpyAsync(bval.data(), c->bvalue.p, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPX(hipStreamSynchronize(c->stream));

    std::vector<int32_t> inside;
    inside.reserve(n);
    for (int64_t i = 0; i < n; ++i)
        if (loc[i] == 1) inside.push_back((int32_t)i);
    const int n_in = (int)inside.size();

    c->last_rounds = 0;
    std::vector<double> h_mean(n_in), h_m2(n_in), h_cnt(n_in), h_ss(n_in);
    std::vector<int32_t> h_hits(n_in);
    double walk_s = 0.0, total_walk_s = 0.0;
    int blocks = 0;
    if (n_in > 0) {
        HIPX(upload(c->inside, inside, c->stream));
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
        HIPX(c->counter.reserve(sizeof(unsigned long long)));
        const bool src = eq->source_kind != WOS_SRC_NONE;
        const bool scr = eq->kind == WOS_EQ_SCREENED;
        const bool adaptive = prm->target_se > 0.0;
        const int64_t cap = adaptive ? std::max<int64_t>(prm->max_walks, prm->n_walks) : prm->n_walks;
        // staging: one FP64 value + one int32 per walk, at most 4 GiB a round
        const int64_t budget = std::max<int64_t>((int64_t)(4LL << 30) / 12, 1);
        std::vector<int64_t> done(n_in, 0), want(n_in, std::min<int64_t>(prm->n_walks, cap));
        std::vector<int32_t> pid, wbeg;
        std::vector<int64_t> off;
        std::vector<double> a_cnt(n_in), a_m2(n_in);
        std::vector<int32_t> all_ids(n_in);
        for (int k = 0; k < n_in; ++k) all_ids[k] = k;
        HIPX(upload(c->pid_all, all_ids, c->stream));
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
                HIPX(hipMemcpyAsync(a_cnt.data(), c->count.p, n_in * sizeof(double), hipMemcpyDeviceToHost, c->stream));
                HIPX(hipMemcpyAsync(a_m2.data(), c->m2.p, n_in * sizeof(double), hipMemcpyDeviceToHost, c->stream));
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
            rv.inside = c->inside.as<int32_t>();
            if (uniform) {
                const int64_t cw = std::min<int64_t>(want[0], std::max<int64_t>(1, budget / n_in));
                rv.pid = c->pid_all.as<int32_t>();
                rv.n = n_in;
                rv.cw = (int)cw;
                rv.w0 = (int)done[0];
                rv.total = cw * n_in;
                for (int k = 0; k < n_in; ++k) { done[k] += cw; want[k] -= cw; }
            } else {
                pid.clear(); wbeg.clear(); off.assign(1, 0);
                int64_t tot = 0;
                for (int k = 0; k < n_in && tot < budget; ++k) {
                    if (want[k] == 0) continue;
                    const int64_t cnt = std::min<int64_t>(want[k], budget - tot);
                    pid.push_back(k);
                    wbeg.push_back((int32_t)done[k]);
                    tot += cnt;
                    off.push_back(tot);
                    done[k] += cnt;
                    want[k] -= cnt;
                }
                HIPX(upload(c->r_pid, pid, c->stream));
                HIPX(upload(c->r_wbeg, wbeg, c->stream));
                HIPX(upload(c->r_off, off, c->stream));
                rv.pid = c->r_pid.as<int32_t>();
                rv.wbeg = c->r_wbeg.as<int32_t>();
                rv.off = c->r_off.as<int64_t>();
                rv.n = (int)pid.size();
                rv.total = tot;
            }
            HIPX(c->vals.reserve((size_t)rv.total * sizeof(double)));
            HIPX(c->steps.reserve((size_t)rv.total * sizeof(int32_t)));
            HIPX(hipEventRecord(c->ev[2], c->stream));
            hipError_t we = dim == 2
                ? dispatch_walk<2>(c, ev, src, scr, prm->precision, rv, prm->epsilon, prm->max_steps, &blocks)
                : dispatch_walk<3>(c, ev, src, scr, prm->precision, rv, prm->epsilon, prm->max_steps, &blocks);
            HIPX(we);
            HIPX(hipEventRecord(c->ev[3], c->stream));
            hipLaunchKernelGGL(k_welford, dim3((rv.n + kBlock - 1) / kBlock), dim3(kBlock), 0,
                               c->stream, rv, c->vals.as<double>(), c->steps.as<int32_t>(),
                               c->mean.as<double>(), c->m2.as<double>(), c->count.as<double>(),
                               c->ssum.as<double>(), c->hits.as<int32_t>());
            HIPX(hipGetLastError());
            HIPX(hipEventRecord(c->ev[4], c->stream));
            HIPX(hipEventSynchronize(c->ev[4]));
            walk_s += elapsed(c->ev[2], c->ev[3]);
            total_walk_s += elapsed(c->ev[2], c->ev[4]);
            ++c->last_rounds;
        }
        HIPX(hipMemcpyAsync(h_mean.data(), c->mean.p, n_in * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPX(hipMemcpyAsync(h_m2.data(), c->m2.p, n_in * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPX(hipMemcpyAsync(h_cnt.data(), c->count.p, n_in * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPX(hipMemcpyAsync(h_ss.data(), c->ssum.p, n_in * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPX(hipMemcpyAsync(h_hits.data(), c->hits.p, n_in * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
        HIPX(hipStreamSynchronize(c->stream));
    }
    const double classify_s = elapsed(c->ev[0], c->ev[1]);

    // per-point result policy (runner.hpp:195-235) and the summary sums in
    // sweep order (runner.hpp:224-228)
    wos_run_summary S{};
    int ii = 0;
    int64_t total_steps = 0;
    for (int64_t i = 0; i < n; ++i) {
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
            total_steps += (int64_t)h_ss[ii];
            ++ii;
            S.inside_count++;
            S.max_steps_hits += hits;
            S.sum_variance += var;
            S.sum_standard_error += se;
            S.sum_mean_steps += ms;
        }
        if (loc[i] == 3) S.indeterminate_points++;
        S.ambiguous_ray_retries += amb[i];
        out->mean[i] = mu;
        if (out->variance) out->variance[i] = var;
        if (out->standard_error) out->standard_error[i] = se;
        if (out->mean_steps) out->mean_steps[i] = ms;
        if (out->location) out->location[i] = loc[i];
        if (out->max_steps_hits) out->max_steps_hits[i] = hits;
        if (out->walks) out->walks[i] = walks;
        if (out->ambiguous_ray_retries) out->ambiguous_ray_retries[i] = amb[i];
    }
    S.wos_seconds = classify_s + total_walk_s;
    S.walk_kernel_seconds = walk_s;
    S.total_walks = 0;
    for (int k = 0; k < n_in; ++k) S.total_walks += (int64_t)h_cnt[k];
    S.rounds = c->last_rounds;
    S.total_steps = total_steps;
    for (int k = 0; k < 6; ++k) S.grid_bounds[k] = gb[k];
    S.boundary_tolerance = tol;
    S.launch_blocks = blocks;
    S.launch_threads = prm->precision == WOS_FP32_PERSISTENT ? c->last_threads : kBlock;
    S.lds_cached_nodes = prm->precision == WOS_FP32_PERSISTENT ? c->last_cached : 0;
    if (summary) *summary = S;
    return WOS_OK;
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
    return c.release();
}

void wos_ctx_destroy(wos_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto &e : c->ev) if (e) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

int wos_ctx_set_mesh(wos_ctx *c, int dim, const double *verts, int64_t n_verts,
                     const int32_t *prims, int64_t n_prims, const int32_t *boundary_ids) {
    if (!c) return fail(WOS_ERR_INPUT, "context is NULL");
    if (dim != 2 && dim != 3) return fail(WOS_ERR_INPUT, "dim must be 2 or 3");
    if (!verts || !prims || n_verts <= 0 || n_prims <= 0)
        return fail(WOS_ERR_MESH, "mesh has no vertices or primitives");
    HostMesh m;
    m.dim = dim;
    m.verts.assign(verts, verts + n_verts * dim);
    m.prims.assign(prims, prims + n_prims * dim);
    m.bids.assign(n_prims, 0);
    if (boundary_ids && dim == 2) m.bids.assign(boundary_ids, boundary_ids + n_prims);
    for (double v : m.verts)
        if (!std::isfinite(v)) return fail(WOS_ERR_MESH, "vertex coordinates must be finite");
    for (int32_t i : m.prims)
        if (i < 0 || i >= n_verts) return fail(WOS_ERR_MESH, "vertex index out of range");
    std::string err;
    HostBVH b = build_reference_bvh(m, &err);
    if (b.nodes.empty()) return fail(WOS_ERR_MESH, err);
    if (b.depth > 60) return fail(WOS_ERR_MESH, "BVH deeper than the 60-level traversal stack");
    DeviceLayouts L = make_layouts(m, b);
    HIPX(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    if (dim == 2) {
        HIPX(upload(c->nd2, L.nd2, s));
        HIPX(upload(c->segd, L.segd, s));
        HIPX(upload(c->nf2, L.nf2, s));
        HIPX(upload(c->segf, L.segf, s));
        HIPX(upload(c->seg_inv, L.seg_inv_vv, s));
    } else {
        HIPX(upload(c->nd3, L.nd3, s));
        HIPX(upload(c->trid, L.trid, s));
        HIPX(upload(c->nf3, L.nf3, s));
        HIPX(upload(c->trif, L.trif, s));
    }
    HIPX(upload(c->slot_prim, b.slots, s));
    HIPX(upload(c->prim_bid, m.bids, s));
    HIPX(upload(c->prims, m.prims, s));
    HIPX(upload(c->verts, m.verts, s));
    MeshView v{};
    v.dim = dim;
    v.n_prims = (int)n_prims;
    v.n_nodes = (int)b.nodes.size();
    v.depth = b.depth;
    v.nd2 = c->nd2.as<NodeD2>(); v.nd3 = c->nd3.as<NodeD3>();
    v.segd = c->segd.as<SegD>(); v.trid = c->trid.as<TriD>();
    v.slot_prim = c->slot_prim.as<int32_t>(); v.prim_bid = c->prim_bid.as<int32_t>();
    v.prims = c->prims.as<int32_t>(); v.verts = c->verts.as<double>();
    v.nf2 = c->nf2.as<NodeF2>(); v.nf3 = c->nf3.as<NodeF3>();
    v.segf = c->segf.as<SegF>(); v.seg_inv_vv = c->seg_inv.as<float>(); v.trif = c->trif.as<TriF>();
    v.root_ref = L.root_ref;
    // conservative FP32 radius margin: 32 ulp of the largest coordinate
    v.shrink_abs = (float)(32.0 * std::ldexp(1.0, -24) * L.coord_scale);
    c->view = v;
    c->mesh = std::move(m);
    c->bvh = std::move(b);
    c->L = std::move(L);
    c->has_mesh = true;
    return WOS_OK;
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
    c->tune = *t;
    return WOS_OK;
}

int wos_ctx_get_tuning(const wos_ctx *c, wos_tuning *t) {
    if (!c || !t) return fail(WOS_ERR_INPUT, "context or tuning is NULL");
    *t = c->tune;
    return WOS_OK;
}

int wos_solve_grid(wos_ctx *c, int nx, int ny, int nz, const wos_eq_desc *eq,
                   const wos_run_params *params, wos_point_stats *out, wos_run_summary *summary) {
    if (!c || !c->has_mesh) return fail(c ? WOS_ERR_STATE : WOS_ERR_INPUT, "no mesh set on this context");
    if (nx < 1 || ny < 1 || nz < 1) return fail(WOS_ERR_INPUT, "grid sizes must be >= 1");
    if (c->mesh.dim == 2) nz = 1;  // runner.hpp:83
    GridSpec g{};
    g.nx = nx; g.ny = ny; g.nz = nz;
    return solve(c, &g, nullptr, (int64_t)nx * ny * nz, eq, params, out, summary);
}

int wos_solve_points(wos_ctx *c, const double *points, int64_t n, const wos_eq_desc *eq,
                     const wos_run_params *params, wos_point_stats *out,
                     wos_run_summary *summary) {
    if (!points) return fail(WOS_ERR_INPUT, "points is NULL");
    return solve(c, nullptr, points, n, eq, params, out, summary);
}

}  // extern "C"
