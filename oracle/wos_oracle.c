/*
 * wos_oracle.c -- CPU restatement of the reference walk-on-spheres path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline); never linked
 * into the product.  See wos_oracle.h for what pins each function.
 *
 * Reference paths below are relative to the reference repository.
 */
#define _GNU_SOURCE
#include "wos_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

#define ORC_PI 3.141592653589793238462643383279502884

/* ===================================================================== */
/* random numbers                                                         */
/* ===================================================================== */

/* include/wos/hash.hpp:7-12 */
uint64_t orc_splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

/* src/prng.cpp:6-15: four splitmix64 words at seed + k*golden, zero guard */
void orc_xoshiro_seed(uint64_t s[4], uint64_t seed) {
    s[0] = orc_splitmix64(seed);
    s[1] = orc_splitmix64(seed + 0x9E3779B97F4A7C15ULL);
    s[2] = orc_splitmix64(seed + 0x3C6EF372FE94F82AULL);
    s[3] = orc_splitmix64(seed + 0xDAA66D2C7DDF743FULL);
    if (!(s[0] | s[1] | s[2] | s[3])) s[0] = 1;
}

static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

/* include/wos/prng.hpp:12-22 (xoshiro256++, Blackman & Vigna) */
uint64_t orc_xoshiro_next(uint64_t s[4]) {
    const uint64_t result = rotl64(s[0] + s[3], 23) + s[0];
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return result;
}

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  No reference
 * counterpart: the counter stream that replaces per-point xoshiro walk
 * streams on the GPU.  Pinned by the Random123 known-answer vectors. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* include/wos/runner.hpp:163-168 */
uint64_t orc_point_seed(uint64_t seed, uint64_t linear_index) {
    return orc_splitmix64(seed ^ orc_splitmix64(linear_index));
}

static const uint64_t INSIDE_STREAM = 0x243F6A8885A308D3ULL;  /* runner.hpp:182 */
static const uint64_t WOS_STREAM = 0x13198A2E03707344ULL;     /* runner.hpp:208 */
static const uint64_t SOURCE_STREAM = 0xA4093822299F31D0ULL;  /* runner.hpp:211 */

/* A draw stream: either a reference xoshiro256++ stream (prng.hpp) or the
 * counter stream of one walk step.  Counter layout (key = the point's walk
 * stream seed): uniform k = 0, 1, 2, ... of step s of walk w takes its top
 * 32 bits from word k mod 4 of Philox(ctr = {w, s, k / 4, 0}) and its next
 * 21 bits from the top of word k mod 4 of Philox(ctr = {w, s, k / 4, 1}).
 * Slots 0 .. dim-2 are the step direction, the source term draws from slot
 * dim-1 on; a uniform ball point draws its radius first. */
typedef struct orc_rng {
    int mode;
    uint64_t *xo;
    uint32_t key[2];
    uint32_t walk, step, next, held;  /* held: block index + 1 in hi/lo (0: none) */
    uint32_t hi[4], lo[4];
} orc_rng;

/* the 53 bits of the next counter-stream uniform */
static uint64_t rng_bits53(orc_rng *r) {
    const uint32_t k = r->next++, blk = k >> 2;
    if (r->held != blk + 1) {
        uint32_t ctr[4] = {r->walk, r->step, blk, 0};
        orc_philox4x32_10(ctr, r->key, r->hi);
        ctr[3] = 1;
        orc_philox4x32_10(ctr, r->key, r->lo);
        r->held = blk + 1;
    }
    return ((uint64_t)r->hi[k & 3] << 21) | (r->lo[k & 3] >> 11);
}

/* include/wos/prng.hpp:25-27; the counter stream: k 2^-53 of its 53 bits */
static double rng_unit(orc_rng *r) {
    if (r->mode == ORC_RNG_REFERENCE)
        return (double)(orc_xoshiro_next(r->xo) >> 11) * (1.0 / (double)(1ULL << 53));
    return (double)rng_bits53(r) * (1.0 / (double)(1ULL << 53));
}

/* include/wos/prng.hpp:31-37 (reference: reject zero).  The counter stream
 * maps to the cell midpoint instead, (2j+1) 2^-53 with j its top 52 bits. */
static double rng_unit_open(orc_rng *r) {
    if (r->mode == ORC_RNG_REFERENCE) {
        double v;
        do { v = rng_unit(r); } while (v == 0.0);
        return v;
    }
    return (double)(rng_bits53(r) | 1ULL) * (1.0 / (double)(1ULL << 53));
}

/* counter stream of step `step` from slot `first` on */
static void rng_begin_step(orc_rng *r, uint32_t step, uint32_t first) {
    r->step = step;
    r->next = first;
    r->held = 0;
}

/* The first n uniforms of a counter step stream from slot `first` on (FP64
 * values; exported for the tests of the stream layout). */
void orc_counter_draws(uint64_t stream_seed, uint32_t walk, uint32_t step, uint32_t first,
                       int n, double *out) {
    orc_rng r;
    memset(&r, 0, sizeof r);
    r.mode = ORC_RNG_COUNTER;
    r.key[0] = (uint32_t)stream_seed;
    r.key[1] = (uint32_t)(stream_seed >> 32);
    r.walk = walk;
    rng_begin_step(&r, step, first);
    for (int i = 0; i < n; ++i) out[i] = rng_unit(&r);
}

/* ===================================================================== */
/* special functions                                                      */
/* ===================================================================== */

/* Large-x series of e^-x I0(x) and e^x K0(x) as in include/wos/bessel.hpp:16-37
 * (the standard order-zero Hankel coefficients, stopped at the smallest term). */
static double asymptotic_sum(double x, int alternating) {
    double sum = 1.0, term = 1.0, previous = INFINITY;
    for (int k = 1; k <= 24; ++k) {
        const double odd = (double)(2 * k - 1);
        double factor = odd * odd / (8.0 * x * (double)k);
        if (alternating) factor = -factor;
        term *= factor;
        const double magnitude = fabs(term);
        if (magnitude > previous) break;
        sum += term;
        if (magnitude <= DBL_EPSILON * fabs(sum)) break;
        previous = magnitude;
    }
    return sum;
}

/* I0 power series sum_k (x^2/4)^k / (k!)^2 (A&S 9.6.12); all terms positive. */
static double i0_series(double x) {
    const double q = 0.25 * x * x;
    double term = 1.0, sum = 1.0;
    for (int k = 1; k < 500; ++k) {
        term *= q / ((double)k * (double)k);
        sum += term;
        if (term <= 1e-17 * sum) break;
    }
    return sum;
}

/* e^-x I0(x).  Reference (bessel.hpp:40-48) calls libstdc++ cyl_bessel_i
 * below x = 50; this restatement uses the power series there instead
 * (parity against the reference at that boundary is to ~1e-15 relative,
 * pinned against scipy.special.i0e). */
double orc_i0_scaled(double x) {
    x = fabs(x);
    if (x < 50.0) return exp(-x) * i0_series(x);
    return asymptotic_sum(x, 0) / sqrt(2.0 * ORC_PI * x);
}

/* e^x K0(x).  Reference (bessel.hpp:50-57) calls libstdc++ cyl_bessel_k below
 * x = 50.  Here: x <= 2 the series K0 = -(ln(x/2)+gamma) I0 + sum (x^2/4)^k
 * H_k/(k!)^2 (A&S 9.6.13); 2 < x < 50 the trapezoidal rule on
 * e^x K0(x) = int_0^inf exp(-2x sinh^2(t/2)) dt (A&S 9.6.24), which converges
 * geometrically in the step for this analytic integrand. */
double orc_k0_scaled(double x) {
    if (x < 0.0) return NAN;
    if (x == 0.0) return INFINITY;
    if (x <= 2.0) {
        const double q = 0.25 * x * x;
        double term = 1.0, hk = 0.0, sum = 0.0, i0 = 1.0;
        for (int k = 1; k < 200; ++k) {
            term *= q / ((double)k * (double)k);
            hk += 1.0 / (double)k;
            i0 += term;
            sum += term * hk;
            if (term <= 1e-18) break;
        }
        const double k0 = -(log(0.5 * x) + 0.57721566490153286060651209) * i0 + sum;
        return exp(x) * k0;
    }
    if (x < 50.0) {
        /* the integrand is about exp(-x t^2 / 2): keep the step below
         * 0.5 / sqrt(x) so the trapezoid's aliasing error stays < e^-70 */
        const double h = fmin(0.125, 0.5 / sqrt(x));
        double sum = 0.5;  /* t = 0 end point, weight 1/2 */
        for (int i = 1; i < 4000; ++i) {
            const double sh = sinh(0.5 * h * (double)i);
            const double v = exp(-2.0 * x * sh * sh);
            sum += v;
            if (v < 1e-18) break;
        }
        return h * sum;
    }
    return sqrt(ORC_PI / (2.0 * x)) * asymptotic_sum(x, 1);
}

/* include/wos/bessel.hpp:60-68 */
double orc_log_i0(double x) {
    x = fabs(x);
    return x + log(orc_i0_scaled(x));
}
double orc_log_k0(double x) {
    if (x < 0.0) return NAN;
    if (x == 0.0) return INFINITY;
    return -x + log(orc_k0_scaled(x));
}

/* include/wos/poisson.hpp:12-18 */
double orc_poisson_green_2d(double R, double r) {
    if (!(r > 0.0) || r > R || R <= 0.0) return NAN;
    if (r == R) return 0.0;
    return log(R / r) / (2.0 * ORC_PI);
}
/* include/wos/poisson.hpp:32-38 */
double orc_poisson_green_3d(double R, double r) {
    if (!(r > 0.0) || r > R || R <= 0.0) return NAN;
    if (r == R) return 0.0;
    return (R - r) / (4.0 * ORC_PI * r * R);
}
/* include/wos/screened.hpp:14-22 */
double orc_screened_weight_2d(double alpha, double R) {
    const double x = alpha * R;
    if (x == 0.0) return 1.0;
    if (x < 1e-3) {
        const double x2 = x * x;
        return 1.0 - x2 / 4.0 + 3.0 * x2 * x2 / 64.0;
    }
    return exp(-orc_log_i0(x));
}
/* include/wos/screened.hpp:24-41 */
double orc_screened_green_2d(double alpha, double R, double r) {
    if (!(r > 0.0) || r > R || R <= 0.0) return NAN;
    if (r == R) return 0.0;
    const double x = alpha * r, X = alpha * R;
    if (X < 1e-5) return orc_poisson_green_2d(R, r);
    const double log_a = orc_log_k0(x);
    const double log_b = orc_log_i0(x) + orc_log_k0(X) - orc_log_i0(X);
    double delta = log_b - log_a;
    if (delta > 0.0) delta = 0.0;
    return exp(log_a) * (-expm1(delta)) / (2.0 * ORC_PI);
}
/* include/wos/screened.hpp:43-52 */
double orc_screened_mass_2d(double alpha, double R) {
    const double x = alpha * R;
    if (x < 1e-3) {
        const double x2 = x * x;
        return R * R * (1.0 / 4.0 - 3.0 * x2 / 64.0 + 19.0 * x2 * x2 / 2304.0);
    }
    return -expm1(-orc_log_i0(x)) / (alpha * alpha);
}
/* include/wos/screened.hpp:54-62 */
double orc_screened_weight_3d(double alpha, double R) {
    const double x = alpha * R;
    if (x == 0.0) return 1.0;
    if (x < 1e-3) {
        const double x2 = x * x;
        return 1.0 - x2 / 6.0 + 7.0 * x2 * x2 / 360.0;
    }
    return 2.0 * x * exp(-x) / (-expm1(-2.0 * x));
}
/* include/wos/screened.hpp:64-80 */
double orc_screened_green_3d(double alpha, double R, double r) {
    if (!(r > 0.0) || r > R || R <= 0.0) return NAN;
    if (r == R) return 0.0;
    const double x = alpha * r, X = alpha * R;
    if (X < 1e-5) return orc_poisson_green_3d(R, r);
    const double ratio = exp(-x) * (-expm1(-2.0 * (X - x))) / (-expm1(-2.0 * X));
    return ratio / (4.0 * ORC_PI * r);
}
/* include/wos/fastmath.hpp:15-29: the series of J0 stopped at the first term
 * below 1e-8 (the reference's own truncation; pinned against scipy's j0 to
 * that accuracy) */
double orc_bessel_j0_series(double x) {
    const double q = 0.25 * x * x;
    double term = 1.0, sum = 1.0;
    for (int k = 1; k < 200; ++k) {
        term *= q / (double)(k * k);
        if (term < 1e-8) return sum;
        sum += (k % 2 == 0 ? term : -term);
    }
    return sum;
}
/* src/equations/helmholtz.cpp:36-41 (2D) and :70-75 (3D) */
double orc_helmholtz_weight_2d(double k, double R) {
    const double prod = R * k;
    if (prod < 1e-8) return 1.0;
    return 1.0 / orc_bessel_j0_series(prod);
}
double orc_helmholtz_weight_3d(double k, double R) {
    const double prod = R * k;
    if (prod < 1e-8) return 1.0;
    return prod / sin(prod);
}
/* include/wos/screened.hpp:82-91 */
double orc_screened_mass_3d(double alpha, double R) {
    const double x = alpha * R;
    if (x < 1e-3) {
        const double x2 = x * x;
        return R * R * (1.0 / 6.0 - 7.0 * x2 / 360.0 + 31.0 * x2 * x2 / 15120.0);
    }
    return (1.0 - orc_screened_weight_3d(alpha, R)) / (alpha * alpha);
}

static double clamp_probability(double v) {           /* screened.hpp:93-95 */
    return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
}

/* include/wos/screened.hpp:97-101 */
static double normal_draw(orc_rng *rng) {
    const double radius = sqrt(-2.0 * log(rng_unit_open(rng)));
    const double angle = 2.0 * ORC_PI * rng_unit(rng);
    return radius * cos(angle);
}

/* include/wos/poisson.hpp:27-30 */
static double poisson_radius_2d(double R, orc_rng *rng) {
    const double u = rng_unit_open(rng);
    const double v = rng_unit_open(rng);
    return R * sqrt(u * v);
}
/* include/wos/poisson.hpp:46-50 */
static double poisson_radius_3d(double R, orc_rng *rng) {
    const double u1 = rng_unit_open(rng), u2 = rng_unit_open(rng);
    const double a = -log(u1 * u2);
    const double v1 = rng_unit_open(rng), v2 = rng_unit_open(rng);
    const double b = -log(v1 * v2);
    return R * a / (a + b);
}

/* include/wos/screened.hpp:105-129 */
static double screened_radius_2d(double alpha, double R, orc_rng *rng) {
    const double X = alpha * R;
    for (;;) {
        double r, acceptance;
        if (X < 2.0) {
            r = poisson_radius_2d(R, rng);
            const double proposal = orc_poisson_green_2d(R, r);
            acceptance = orc_screened_green_2d(alpha, R, r) / proposal;
        } else {
            const double time = -log(rng_unit_open(rng)) / (alpha * alpha);
            r = sqrt(-4.0 * time * log(rng_unit_open(rng)));
            if (r >= R) continue;
            const double x = alpha * r;
            const double log_ratio = orc_log_i0(x) + orc_log_k0(X) - orc_log_i0(X) - orc_log_k0(x);
            acceptance = -expm1(log_ratio < 0.0 ? log_ratio : 0.0);
        }
        if (rng_unit(rng) < clamp_probability(acceptance)) return r;
    }
}

/* include/wos/screened.hpp:131-156 */
static double screened_radius_3d(double alpha, double R, orc_rng *rng) {
    const double X = alpha * R;
    for (;;) {
        double r, acceptance;
        if (X < 2.0) {
            r = poisson_radius_3d(R, rng);
            const double proposal = orc_poisson_green_3d(R, r);
            acceptance = orc_screened_green_3d(alpha, R, r) / proposal;
        } else {
            const double time = -log(rng_unit_open(rng)) / (alpha * alpha);
            const double scale = sqrt(2.0 * time);
            const double x0 = scale * normal_draw(rng);
            const double x1 = scale * normal_draw(rng);
            const double x2 = scale * normal_draw(rng);
            r = sqrt(x0 * x0 + x1 * x1 + x2 * x2);
            if (r >= R) continue;
            const double x = alpha * r;
            acceptance = (-expm1(-2.0 * (X - x))) / (-expm1(-2.0 * X));
        }
        if (rng_unit(rng) < clamp_probability(acceptance)) return r;
    }
}

/* ===================================================================== */
/* primitives                                                             */
/* ===================================================================== */

/* src/npq.cpp:8-29; also returns the clamped segment parameter */
double orc_npq_seg(const double s0[2], const double s1[2], const double p[2],
                   double nearest[2], double *par_out) {
    const double vx = s1[0] - s0[0], vy = s1[1] - s0[1];
    const double wx = p[0] - s0[0], wy = p[1] - s0[1];
    const double vv = vx * vx + vy * vy;
    const double wv = wx * vx + wy * vy;
    double par = (vv > 0.0) ? wv / vv : 0.0;
    if (par < 0.0) par = 0.0;
    if (par > 1.0) par = 1.0;
    nearest[0] = s0[0] + par * vx;
    nearest[1] = s0[1] + par * vy;
    const double dx = p[0] - nearest[0], dy = p[1] - nearest[1];
    if (par_out) *par_out = par;
    return dx * dx + dy * dy;
}

static inline double dot3(const double *a, const double *b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static inline void cross3(const double *a, const double *b, double *o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

/* src/npq.cpp:52-189 with the EberlyCache fill of src/bvh.cpp:190-205
 * (edge1 = v1-v0, edge2 = v2-v0, a,b,c, det = ac-b^2, denom = a-2b+c).
 * Returns the squared distance; st receives the (s, t) of the nearest point
 * v0 + s*edge1 + t*edge2. */
double orc_eberly(const double p[3], const double v0[3], const double v1[3],
                  const double v2[3], double nearest[3], double st[2]) {
    const double e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
    const double e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
    const double a = dot3(e1, e1), b = dot3(e1, e2), c = dot3(e2, e2);
    const double det = a * c - b * b;
    const double denom = a - 2 * b + c;
    const double BP[3] = {v0[0] - p[0], v0[1] - p[1], v0[2] - p[2]};
    const double d = dot3(e1, BP), e = dot3(e2, BP);
    double s = b * e - c * d;
    double t = b * d - a * e;
    if (s + t <= det) {
        if (s < 0) {
            if (t < 0) {                      /* region 4 */
                if (d < 0) {
                    t = 0;
                    if (-d >= a) s = 1; else s = -d / a;
                } else {
                    s = 0;
                    if (e >= 0) t = 0; else if (-e >= c) t = 1; else t = -e / c;
                }
            } else {                          /* region 3 */
                s = 0;
                if (e >= 0) t = 0; else if (-e >= c) t = 1; else t = -e / c;
            }
        } else if (t < 0) {                   /* region 5 */
            t = 0;
            if (d >= 0) s = 0; else if (-d >= a) s = 1; else s = -d / a;
        } else {                              /* region 0 */
            s /= det;
            t /= det;
        }
    } else {
        if (s < 0) {                          /* region 2 */
            const double bd = b + d, ce = c + e;
            if (ce > bd) {
                const double numer = ce - bd;
                if (numer >= denom) s = 1; else s = numer / denom;
                t = 1 - s;
            } else {
                s = 0;
                if (ce <= 0) t = 1; else if (e >= 0) t = 0; else t = -e / c;
            }
        } else if (t < 0) {                   /* region 6 */
            const double be = b + e, ad = a + d;
            if (ad > be) {
                const double numer = ad - be;
                if (numer >= denom) t = 1; else t = numer / denom;
                s = 1 - t;
            } else {
                t = 0;
                if (ad <= 0) s = 1; else if (d >= 0) s = 0; else s = -d / a;
            }
        } else {                              /* region 1 */
            const double numer = (c + e) - (b + d);
            if (numer <= 0) s = 0;
            else if (numer >= denom) s = 1;
            else s = numer / denom;
            t = 1 - s;
        }
    }
    nearest[0] = v0[0] + s * e1[0] + t * e2[0];
    nearest[1] = v0[1] + s * e1[1] + t * e2[1];
    nearest[2] = v0[2] + s * e1[2] + t * e2[2];
    const double dv[3] = {nearest[0] - p[0], nearest[1] - p[1], nearest[2] - p[2]};
    if (st) { st[0] = s; st[1] = t; }
    return dot3(dv, dv);
}

#define INTERSECTION_EPSILON 1e-9   /* src/inside.cpp:14 */

/* src/inside.cpp:140-182; returns 0 miss, 1 hit, 2 ambiguous */
int orc_moller_trumbore(const double p[3], const double ray[3], const double a[3],
                        const double b[3], const double c[3]) {
    const double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    const double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    const double s[3] = {p[0] - a[0], p[1] - a[1], p[2] - a[2]};
    double rxe2[3], normal[3], sxe1[3];
    cross3(ray, e2, rxe2);
    const double det = dot3(e1, rxe2);
    cross3(e1, e2, normal);
    const double normal_length = sqrt(dot3(normal, normal));
    const double det_tolerance = INTERSECTION_EPSILON * normal_length;
    if (fabs(det) <= det_tolerance) {
        if (normal_length == 0.0) return 0;
        double edge_scale = 1.0;
        const double l1 = sqrt(dot3(e1, e1)), l2 = sqrt(dot3(e2, e2));
        if (l1 > edge_scale) edge_scale = l1;
        if (l2 > edge_scale) edge_scale = l2;
        const double plane_distance = fabs(dot3(normal, s)) / normal_length;
        return plane_distance <= INTERSECTION_EPSILON * edge_scale ? 2 : 0;
    }
    const double u = dot3(s, rxe2) / det;
    if (u < -INTERSECTION_EPSILON || u - 1.0 > INTERSECTION_EPSILON) return 0;
    cross3(s, e1, sxe1);
    const double v = dot3(ray, sxe1) / det;
    if (v < -INTERSECTION_EPSILON || u + v - 1.0 > INTERSECTION_EPSILON) return 0;
    const double t = dot3(e2, sxe1) / det;
    if (t <= INTERSECTION_EPSILON) return 0;
    const double w = 1.0 - u - v;
    if (u <= INTERSECTION_EPSILON || v <= INTERSECTION_EPSILON || w <= INTERSECTION_EPSILON)
        return 2;
    return 1;
}

/* ===================================================================== */
/* mesh + BVH                                                             */
/* ===================================================================== */

typedef struct build_item {
    double bmin[3], bmax[3], centroid[3];
    int prim_idx;
} build_item;

static int g_sort_axis;  /* qsort has no context argument; the build is serial */

/* The reference's partition order (src/bvh.cpp:141-171): by centroid along
 * the axis, ties by primitive index.  The total order makes the two sides of
 * std::nth_element implementation independent; a full sort of the range
 * yields the same sets. */
static int cmp_items(const void *pa, const void *pb) {
    const build_item *a = (const build_item *)pa, *b = (const build_item *)pb;
    const double ca = a->centroid[g_sort_axis], cb = b->centroid[g_sort_axis];
    if (ca < cb) return -1;
    if (ca > cb) return 1;
    return (a->prim_idx > b->prim_idx) - (a->prim_idx < b->prim_idx);
}

/* src/bvh.cpp:174-230 */
static int build_subtree(orc_bvh *bvh, build_item *items, int start, int end, int *node_ptr) {
    const int dim = bvh->dim;
    const int span = end - start;
    orc_node *nd = &bvh->nodes[*node_ptr];
    const int node = (*node_ptr)++;
    /* union_bbox (bvh.cpp:123-138): dmin/dmax of the primitive boxes */
    for (int k = 0; k < 3; ++k) { nd->bmin[k] = items[start].bmin[k]; nd->bmax[k] = items[start].bmax[k]; }
    for (int i = start + 1; i < end; ++i)
        for (int k = 0; k < dim; ++k) {
            /* dmin/dmax of include/wos/fastmath.hpp:10-11 */
            nd->bmin[k] = nd->bmin[k] < items[i].bmin[k] ? nd->bmin[k] : items[i].bmin[k];
            nd->bmax[k] = nd->bmax[k] > items[i].bmax[k] ? nd->bmax[k] : items[i].bmax[k];
        }
    if (span <= 4) {                                   /* LEAF_SIZE (bvh.cpp:15) */
        nd->prim_count = span;
        nd->idx = start;
        for (int i = 0; i < span; ++i) bvh->prims[start + i] = items[start + i].prim_idx;
        return node;
    }
    nd->prim_count = 0;
    const double ex = nd->bmax[0] - nd->bmin[0], ey = nd->bmax[1] - nd->bmin[1];
    int axis;
    if (dim == 2) axis = (ex >= ey) ? 0 : 1;
    else {
        const double ez = nd->bmax[2] - nd->bmin[2];
        axis = (ex >= ey) ? (ex >= ez ? 0 : 2) : (ey >= ez ? 1 : 2);
    }
    const int mid = (start + end) / 2;
    g_sort_axis = axis;
    qsort(items + start, (size_t)span, sizeof(build_item), cmp_items);
    build_subtree(bvh, items, start, mid, node_ptr);
    const int right = build_subtree(bvh, items, mid, end, node_ptr);
    bvh->nodes[node].idx = right;
    return node;
}

static const double *vtx(const orc_bvh *b, int i) { return b->verts + (size_t)i * b->dim; }

/* src/bvh.cpp:234-253 with prim_bbox/centroid of src/mesh.cpp:324-366 */
orc_bvh *orc_bvh_build(int dim, const double *verts, int n_verts, const int *prims,
                       int n_prims, const int *bids) {
    if ((dim != 2 && dim != 3) || n_prims <= 0 || n_verts <= 0) return NULL;
    orc_bvh *b = (orc_bvh *)calloc(1, sizeof(orc_bvh));
    b->dim = dim; b->n_verts = n_verts; b->n_prims = n_prims;
    b->verts = (double *)malloc(sizeof(double) * (size_t)n_verts * dim);
    memcpy(b->verts, verts, sizeof(double) * (size_t)n_verts * dim);
    b->mesh_prims = (int *)malloc(sizeof(int) * (size_t)n_prims * dim);
    memcpy(b->mesh_prims, prims, sizeof(int) * (size_t)n_prims * dim);
    b->bids = (int *)calloc((size_t)n_prims, sizeof(int));
    if (bids) memcpy(b->bids, bids, sizeof(int) * (size_t)n_prims);
    b->nodes = (orc_node *)calloc((size_t)(2 * n_prims - 1), sizeof(orc_node));
    b->prims = (int *)malloc(sizeof(int) * (size_t)n_prims);
    build_item *items = (build_item *)calloc((size_t)n_prims, sizeof(build_item));
    for (int p = 0; p < n_prims; ++p) {
        build_item *it = &items[p];
        it->prim_idx = p;
        const double *v0 = vtx(b, prims[dim * p]), *v1 = vtx(b, prims[dim * p + 1]);
        if (dim == 2) {
            for (int k = 0; k < 2; ++k) {
                it->bmin[k] = fmin(v0[k], v1[k]);
                it->bmax[k] = fmax(v0[k], v1[k]);
                it->centroid[k] = (v0[k] + v1[k]) / 2.0;
            }
        } else {
            const double *v2 = vtx(b, prims[dim * p + 2]);
            for (int k = 0; k < 3; ++k) {
                it->bmin[k] = fmin(fmin(v0[k], v1[k]), v2[k]);
                it->bmax[k] = fmax(fmax(v0[k], v1[k]), v2[k]);
                it->centroid[k] = (v0[k] + v1[k] + v2[k]) / 3.0;
            }
        }
    }
    int node_ptr = 0;
    build_subtree(b, items, 0, n_prims, &node_ptr);
    b->n_nodes = node_ptr;
    free(items);
    return b;
}

void orc_bvh_free(orc_bvh *b) {
    if (!b) return;
    free(b->verts); free(b->mesh_prims); free(b->bids); free(b->nodes); free(b->prims);
    free(b);
}

int orc_bvh_n_nodes(const orc_bvh *b) { return b->n_nodes; }

void orc_bvh_export(const orc_bvh *b, int *prim_count, int *idx, double *bmin,
                    double *bmax, int *leaf_prims) {
    for (int i = 0; i < b->n_nodes; ++i) {
        if (prim_count) prim_count[i] = b->nodes[i].prim_count;
        if (idx) idx[i] = b->nodes[i].idx;
        for (int k = 0; k < 3; ++k) {
            if (bmin) bmin[3 * i + k] = b->nodes[i].bmin[k];
            if (bmax) bmax[3 * i + k] = b->nodes[i].bmax[k];
        }
    }
    if (leaf_prims) memcpy(leaf_prims, b->prims, sizeof(int) * (size_t)b->n_prims);
}

/* src/mesh.cpp:296-322 */
void orc_mesh_bbox(const orc_bvh *b, double g[6]) {
    const double *p0 = vtx(b, 0);
    double lo[3] = {p0[0], p0[1], b->dim == 3 ? p0[2] : 0.0};
    double hi[3] = {lo[0], lo[1], lo[2]};
    for (int i = 1; i < b->n_verts; ++i) {
        const double *v = vtx(b, i);
        for (int k = 0; k < b->dim; ++k) {
            if (v[k] < lo[k]) lo[k] = v[k];
            else if (v[k] > hi[k]) hi[k] = v[k];
        }
    }
    g[0] = lo[0]; g[1] = hi[0]; g[2] = lo[1]; g[3] = hi[1]; g[4] = lo[2]; g[5] = hi[2];
}

/* src/bvh.cpp:111-121 */
static double aabb_d_sq(const orc_node *n, const double *p, int dim) {
    double acc = 0.0;
    for (int k = 0; k < dim; ++k) {
        const double lo = n->bmin[k] - p[k], hi = p[k] - n->bmax[k];
        double d = lo > hi ? lo : hi;
        d = 0.0 > d ? 0.0 : d;
        acc = (k == 0) ? d * d : acc + d * d;
    }
    return acc;
}

static double prim_npq(const orc_bvh *b, int s, const double *p, double *np, double *st) {
    if (b->dim == 2) {
        return orc_npq_seg(vtx(b, b->mesh_prims[2 * s]), vtx(b, b->mesh_prims[2 * s + 1]), p,
                           np, &st[0]);
    }
    return orc_eberly(p, vtx(b, b->mesh_prims[3 * s]), vtx(b, b->mesh_prims[3 * s + 1]),
                      vtx(b, b->mesh_prims[3 * s + 2]), np, st);
}

/* src/bvh.cpp:295-387: depth-first, nearer child popped first, strict <
 * keeps the first primitive at the minimum distance.  Counts node records
 * read (the popped node plus both child boxes of an internal node) and
 * primitive tests, the algorithmic work of one query. */
void orc_bvh_npq(const orc_bvh *b, const double *p, orc_npq *out, int64_t *node_reads,
                 int64_t *prim_tests) {
    double best = INFINITY;
    int stack[64], sp = 0;
    stack[sp++] = 0;
    out->prim = -1; out->bid = -1; out->s = 0; out->t = 0;
    out->point[0] = out->point[1] = out->point[2] = 0.0;
    int64_t nr = 0, pt = 0;
    while (sp > 0) {
        const int curr = stack[--sp];
        const orc_node *n = &b->nodes[curr];
        ++nr;
        if (aabb_d_sq(n, p, b->dim) >= best) continue;
        if (n->prim_count > 0) {
            for (int i = 0; i < n->prim_count; ++i) {
                const int s = b->prims[n->idx + i];
                double np[3] = {0, 0, 0}, st[2] = {0, 0};
                const double d = prim_npq(b, s, p, np, st);
                ++pt;
                if (d < best) {
                    best = d;
                    memcpy(out->point, np, sizeof(np));
                    out->prim = s;
                    out->bid = b->dim == 2 ? b->bids[s] : -1;
                    out->s = st[0]; out->t = b->dim == 3 ? st[1] : 0.0;
                }
            }
        } else {
            const int l = curr + 1, r = n->idx;
            const double ld = aabb_d_sq(&b->nodes[l], p, b->dim);
            const double rd = aabb_d_sq(&b->nodes[r], p, b->dim);
            nr += 2;
            if (ld <= rd) { stack[sp++] = r; stack[sp++] = l; }
            else { stack[sp++] = l; stack[sp++] = r; }
        }
    }
    out->distance = sqrt(best);
    if (node_reads) *node_reads += nr;
    if (prim_tests) *prim_tests += pt;
}

/* src/npq.cpp:32-49, :192-220 (brute force over every primitive) */
void orc_npq_brute(const orc_bvh *b, const double *p, orc_npq *out) {
    double best = INFINITY;
    out->prim = -1; out->bid = -1; out->s = out->t = 0;
    for (int s = 0; s < b->n_prims; ++s) {
        double np[3] = {0, 0, 0}, st[2] = {0, 0};
        const double d = prim_npq(b, s, p, np, st);
        if (d < best) {
            best = d;
            memcpy(out->point, np, sizeof(np));
            out->prim = s;
            out->bid = b->dim == 2 ? b->bids[s] : -1;
            out->s = st[0]; out->t = b->dim == 3 ? st[1] : 0.0;
        }
    }
    out->distance = sqrt(best);
}

/* ===================================================================== */
/* classification                                                         */
/* ===================================================================== */

/* src/inside.cpp:117-137 (Sunday winding number over every segment) */
int orc_winding_number(const orc_bvh *b, const double p[2]) {
    int wn = 0;
    for (int s = 0; s < b->n_prims; ++s) {
        const double *s0 = vtx(b, b->mesh_prims[2 * s]), *s1 = vtx(b, b->mesh_prims[2 * s + 1]);
        const double area = (s1[0] - s0[0]) * (p[1] - s0[1]) - (p[0] - s0[0]) * (s1[1] - s0[1]);
        if (s0[1] <= p[1]) {
            if (s1[1] > p[1] && area > 0) wn++;
        } else {
            if (s1[1] <= p[1] && area < 0) wn--;
        }
    }
    return wn;
}

/* src/inside.cpp:27-51 (padded slab update) */
static int update_ray_interval(double origin, double direction, double box_min, double box_max,
                               double *t_min, double *t_max) {
    const double ext = fabs(box_max - box_min);
    const double extent_scale = ext > 1.0 ? ext : 1.0;
    double coord = 1.0;
    if (fabs(origin) > coord) coord = fabs(origin);
    if (fabs(box_min) > coord) coord = fabs(box_min);
    if (fabs(box_max) > coord) coord = fabs(box_max);
    const double padding = INTERSECTION_EPSILON * extent_scale + 16.0 * DBL_EPSILON * coord;
    box_min -= padding;
    box_max += padding;
    if (direction == 0.0) return origin >= box_min && origin <= box_max;
    double t0 = (box_min - origin) / direction;
    double t1 = (box_max - origin) / direction;
    if (t0 > t1) { const double tmp = t0; t0 = t1; t1 = tmp; }
    if (t0 > *t_min) *t_min = t0;
    if (t1 < *t_max) *t_max = t1;
    return *t_min <= *t_max;
}

/* src/inside.cpp:53-62 */
static int ray_hits_box(const double *o, const double *ray, const orc_node *n) {
    double tmin = 0.0, tmax = INFINITY;
    return update_ray_interval(o[0], ray[0], n->bmin[0], n->bmax[0], &tmin, &tmax) &&
           update_ray_interval(o[1], ray[1], n->bmin[1], n->bmax[1], &tmin, &tmax) &&
           update_ray_interval(o[2], ray[2], n->bmin[2], n->bmax[2], &tmin, &tmax);
}

/* src/inside.cpp:78-111: every primitive of every leaf whose padded box the
 * ray meets; an ambiguous primitive abandons the ray */
int orc_ray_cast(const orc_bvh *b, const double p[3], const double ray[3], int *ambiguous) {
    int hits = 0, stack[64], sp = 0;
    stack[sp++] = 0;
    *ambiguous = 0;
    while (sp > 0) {
        const int ni = stack[--sp];
        const orc_node *n = &b->nodes[ni];
        if (!ray_hits_box(p, ray, n)) continue;
        if (n->prim_count > 0) {
            for (int i = 0; i < n->prim_count; ++i) {
                const int tri = b->prims[n->idx + i];
                const int h = orc_moller_trumbore(p, ray, vtx(b, b->mesh_prims[3 * tri]),
                                                  vtx(b, b->mesh_prims[3 * tri + 1]),
                                                  vtx(b, b->mesh_prims[3 * tri + 2]));
                if (h == 2) { *ambiguous = 1; return hits; }
                if (h == 1) ++hits;
            }
        } else {
            stack[sp++] = n->idx;
            stack[sp++] = ni + 1;
        }
    }
    return hits;
}

/* src/inside.cpp:18-25 */
static void random_unit_direction(uint64_t xo[4], double out[3]) {
    const double u = (double)(orc_xoshiro_next(xo) >> 11) * (1.0 / (double)(1ULL << 53));
    const double v = (double)(orc_xoshiro_next(xo) >> 11) * (1.0 / (double)(1ULL << 53));
    const double z = 2 * u - 1;
    const double phi = 2 * ORC_PI * v;
    const double r = sqrt(1 - z * z);
    out[0] = r * cos(phi);
    out[1] = r * sin(phi);
    out[2] = z;
}

/* src/inside.cpp:196-200 (2D) and :228-260 (3D ray vote).  Returns the
 * PointLocation code (inside.hpp:8-13). */
int orc_classify(const orc_bvh *b, const double *p, double distance, double tol,
                 int max_ray_attempts, uint64_t inside_seed, int *ambiguous_retries) {
    *ambiguous_retries = 0;
    if (distance <= tol) return 2;
    if (b->dim == 2) return orc_winding_number(b, p) != 0 ? 1 : 0;
    uint64_t xo[4];
    orc_xoshiro_seed(xo, inside_seed);
    int valid = 0, votes = 0;
    for (int attempts = 0; attempts < max_ray_attempts && valid < 3; ++attempts) {
        double ray[3];
        int amb = 0;
        random_unit_direction(xo, ray);
        const int hits = orc_ray_cast(b, p, ray, &amb);
        if (amb) { ++*ambiguous_retries; continue; }
        votes += hits & 1;
        ++valid;
    }
    if (valid < 3) return 3;
    return votes > 1 ? 1 : 0;
}

/* ===================================================================== */
/* equation functors                                                      */
/* ===================================================================== */

/* FieldSampler2D::sample (apps/welding/include/welding/field_sampler.hpp:51-117):
 * grid coordinates clamped to [0, n-1], then bilinear on the cell, or the
 * tensor cubic Lagrange interpolant on the 4 x 4 nodes starting one node
 * below the cell (shifted inside at the edges) when both n >= 4. */
static double clampd(double v, double lo, double hi) { return v < lo ? lo : (hi < v ? hi : v); }

static void cubic_weights(double c, double w[4]) {
    w[0] = -(c - 1.0) * (c - 2.0) * (c - 3.0) / 6.0;
    w[1] = c * (c - 2.0) * (c - 3.0) / 2.0;
    w[2] = -c * (c - 1.0) * (c - 3.0) / 2.0;
    w[3] = c * (c - 1.0) * (c - 2.0) / 6.0;
}

double orc_field_sample(const wos_field_source *f, double x, double y) {
    const int nx = f->nx, ny = f->ny;
    const double *v = f->values;
    const double gx = clampd((x - f->bounds[0]) / (f->bounds[1] - f->bounds[0]) * (double)(nx - 1),
                             0.0, (double)(nx - 1));
    const double gy = clampd((y - f->bounds[2]) / (f->bounds[3] - f->bounds[2]) * (double)(ny - 1),
                             0.0, (double)(ny - 1));
    if (f->interpolation == WOS_INTERP_CUBIC && nx >= 4 && ny >= 4) {
        int ix = (int)floor(gx) - 1, iy = (int)floor(gy) - 1;
        ix = ix < 0 ? 0 : (ix > nx - 4 ? nx - 4 : ix);
        iy = iy < 0 ? 0 : (iy > ny - 4 ? ny - 4 : iy);
        double wx[4], wy[4];
        cubic_weights(gx - (double)ix, wx);
        cubic_weights(gy - (double)iy, wy);
        double value = 0.0;
        for (int di = 0; di < 4; ++di) {
            double row = 0.0;
            for (int dj = 0; dj < 4; ++dj) row += wy[dj] * v[(size_t)(ix + di) * ny + iy + dj];
            value += wx[di] * row;
        }
        return value;
    }
    const int i0 = (int)floor(gx), j0 = (int)floor(gy);
    const int i1 = i0 + 1 < nx - 1 ? i0 + 1 : nx - 1, j1 = j0 + 1 < ny - 1 ? j0 + 1 : ny - 1;
    const double tx = gx - (double)i0, ty = gy - (double)j0;
    const double v00 = v[(size_t)i0 * ny + j0], v10 = v[(size_t)i1 * ny + j0];
    const double v01 = v[(size_t)i0 * ny + j1], v11 = v[(size_t)i1 * ny + j1];
    const double lower = (1.0 - tx) * v00 + tx * v10;
    const double upper = (1.0 - tx) * v01 + tx * v11;
    return (1.0 - ty) * lower + ty * upper;
}

/* the welding sources, f = history + heat (level0_cooling.cpp:194-196,
 * level2_transient_heat.cpp:255-264, level3_moving_heat.cpp:189-194; heat
 * terms as gaussian_heat_source.hpp:38-41, :109-113 and
 * moving_gaussian_heat_source.hpp:75-88) */
double orc_field_source_value(const wos_field_source *f, double x, double y) {
    double history = 0.0;
    if (f->values) {
        const double v = orc_field_sample(f, x, y);
        history = f->history_scale * (f->clamp_nonnegative ? (v < 0.0 ? 0.0 : v) : v);
    }
    double heat = 0.0;
    if (f->heat_scale != 0.0 && f->n_heat_terms > 0) {
        double sum = 0.0;
        for (int c = 0; c < f->n_heat_terms; ++c) {
            const wos_heat_term *h = &f->heat[c];
            const double dx = (x - h->cx) / h->sx, dy = (y - h->cy) / h->sy;
            sum += h->a * exp(-0.5 * (dx * dx + dy * dy)) / h->b;
        }
        heat = f->heat_scale * sum / f->heat_divisor;
    }
    return history + heat;
}

/* source f(x): poisson.cpp:34-41, screened_poisson.cpp:58-61, plus the
 * descriptor kinds the benchmark problems add */
static double eq_source(const wos_eq_desc *eq, const double *x, int dim) {
    switch (eq->source_kind) {
    case WOS_SRC_FIELD: return orc_field_source_value(eq->field_source, x[0], x[1]);
    case WOS_SRC_DISK_SQ: {
        double acc = 0.0;
        for (int k = 0; k < dim; ++k) {
            const double d = x[k] - eq->src_center[k];
            acc = k == 0 ? d * d : acc + d * d;
        }
        return acc <= eq->src_radius * eq->src_radius ? eq->src_value : 0.0;
    }
    case WOS_SRC_DISK_SQRT: {
        double acc = 0.0;
        for (int k = 0; k < dim; ++k) {
            const double d = x[k] - eq->src_center[k];
            acc = k == 0 ? d * d : acc + d * d;
        }
        return sqrt(acc) <= eq->src_radius ? eq->src_value : 0.0;
    }
    case WOS_SRC_CONSTANT: return eq->src_value;
    case WOS_SRC_GAUSSIAN: {
        double acc = 0.0;
        for (int k = 0; k < dim; ++k) {
            const double d = x[k] - eq->src_center[k];
            acc = k == 0 ? d * d : acc + d * d;
        }
        return eq->src_value * exp(-acc / (eq->src_radius * eq->src_radius));
    }
    default: return 0.0;
    }
}

/* 2D prefilter source_may_intersect (poisson.cpp:27-32,
 * screened_poisson.cpp:48-56); other source kinds have global support */
static int eq_source_may_intersect(const wos_eq_desc *eq, const double *c, double R) {
    if (eq->source_kind != WOS_SRC_DISK_SQ && eq->source_kind != WOS_SRC_DISK_SQRT) return 1;
    const double dx = c[0] - eq->src_center[0], dy = c[1] - eq->src_center[1];
    const double reach = R + eq->src_radius;
    return dx * dx + dy * dy <= reach * reach;
}

/* boundary g: laplace.cpp:16-18 and :30-36, poisson.cpp:43-48, plus the
 * descriptor kinds (per-id constants, per-vertex blend, sin 3x cos 2y) */
static double eq_boundary(const wos_eq_desc *eq, const orc_bvh *b, const orc_npq *q) {
    const double *p = q->point;
    switch (eq->boundary_kind) {
    case WOS_BND_X2Y2_ID0:
        return q->bid == 0 ? p[0] * p[0] - p[1] * p[1] : 0.0;
    case WOS_BND_Y5: {
        const double r_sq = p[0] * p[0] + p[1] * p[1];
        const double y50 = p[2] * (8 * p[2] * p[2] * p[2] * p[2] - 40 * p[2] * p[2] * r_sq + 15 * r_sq * r_sq);
        const double y5m3 = (8 * p[2] * p[2] - r_sq) * p[1] * (3 * p[0] * p[0] - p[1] * p[1]);
        return y50 + 4.0 * y5m3;
    }
    case WOS_BND_ID_CONST:
        return (q->bid >= 0 && q->bid < eq->n_bnd_values) ? eq->bnd_values[q->bid] : 0.0;
    case WOS_BND_VERTEX: {
        const int *pr = b->mesh_prims + (size_t)q->prim * b->dim;
        const double *vv = eq->vertex_values;
        if (b->dim == 2) return (1.0 - q->s) * vv[pr[0]] + q->s * vv[pr[1]];
        return (1.0 - q->s - q->t) * vv[pr[0]] + q->s * vv[pr[1]] + q->t * vv[pr[2]];
    }
    case WOS_BND_SIN3X_COS2Y: return sin(3.0 * p[0]) * cos(2.0 * p[1]);
    case WOS_BND_ONE: return 1.0;
    default: return 0.0;
    }
}

static double eq_green(const wos_eq_desc *eq, int dim, double R, double r) {
    if (eq->kind == WOS_EQ_SCREENED)
        return dim == 2 ? orc_screened_green_2d(eq->alpha, R, r) : orc_screened_green_3d(eq->alpha, R, r);
    return dim == 2 ? orc_poisson_green_2d(R, r) : orc_poisson_green_3d(R, r);
}
static double eq_green_mass(const wos_eq_desc *eq, int dim, double R) {
    if (eq->kind == WOS_EQ_SCREENED)
        return dim == 2 ? orc_screened_mass_2d(eq->alpha, R) : orc_screened_mass_3d(eq->alpha, R);
    return dim == 2 ? R * R / 4.0 : R * R / 6.0;
}

/* sample_green of poisson.cpp:54-62/96-107 and screened_poisson.cpp:69-77/116-127 */
static void eq_sample_green(const wos_eq_desc *eq, int dim, const double *c, double R,
                            orc_rng *rng, double *out) {
    if (dim == 2) {
        const double radius = eq->kind == WOS_EQ_SCREENED ? screened_radius_2d(eq->alpha, R, rng)
                                                          : poisson_radius_2d(R, rng);
        const double angle = 2.0 * ORC_PI * rng_unit(rng);
        out[0] = c[0] + radius * cos(angle);
        out[1] = c[1] + radius * sin(angle);
    } else {
        const double radius = eq->kind == WOS_EQ_SCREENED ? screened_radius_3d(eq->alpha, R, rng)
                                                          : poisson_radius_3d(R, rng);
        const double z = 2.0 * rng_unit(rng) - 1.0;
        const double angle = 2.0 * ORC_PI * rng_unit(rng);
        const double xy = sqrt(1.0 - z * z);
        out[0] = c[0] + radius * xy * cos(angle);
        out[1] = c[1] + radius * xy * sin(angle);
        out[2] = c[2] + radius * z;
    }
}

/* src/wos.cpp:33-56 (uniform point in the ball); the counter stream draws
 * the radius first, the reference's stream the direction first */
static void sphere_sample(int dim, const double *c, double R, orc_rng *rng, double *out) {
    const int radius_first = rng->mode != ORC_RNG_REFERENCE;
    if (dim == 2) {
        double theta, r;
        if (radius_first) {
            r = R * sqrt(rng_unit_open(rng));
            theta = rng_unit(rng) * 2 * ORC_PI;
        } else {
            theta = rng_unit(rng) * 2 * ORC_PI;
            r = R * sqrt(rng_unit_open(rng));
        }
        out[0] = c[0] + r * cos(theta);
        out[1] = c[1] + r * sin(theta);
    } else {
        double u, v, w;
        if (radius_first) {
            w = rng_unit_open(rng);
            u = rng_unit(rng);
            v = rng_unit(rng);
        } else {
            u = rng_unit(rng);
            v = rng_unit(rng);
            w = rng_unit_open(rng);
        }
        const double z = 2 * u - 1, phi = 2 * ORC_PI * v;
        const double surf = sqrt(1 - z * z), vol = R * cbrt(w);
        out[0] = c[0] + vol * surf * cos(phi);
        out[1] = c[1] + vol * surf * sin(phi);
        out[2] = c[2] + vol * z;
    }
}

/* include/wos/wos.hpp:21-27 */
static double sphere_volume(int dim, double R) {
    if (dim == 2) return ORC_PI * R * R;
    const double rc = R * R * R;
    return 4.0 * ORC_PI * rc / 3.0;
}

/* src/wos.cpp:9-30 (uniform point on the sphere) */
static void step_on_sphere(int dim, const double *c, double R, orc_rng *rng, double *out) {
    if (dim == 2) {
        const double theta = rng_unit(rng) * 2 * ORC_PI;
        out[0] = c[0] + R * cos(theta);
        out[1] = c[1] + R * sin(theta);
    } else {
        const double u = rng_unit(rng), v = rng_unit(rng);
        const double z = 2 * u - 1, phi = 2 * ORC_PI * v;
        const double r = sqrt(1 - z * z);
        out[0] = c[0] + R * r * cos(phi);
        out[1] = c[1] + R * r * sin(phi);
        out[2] = c[2] + R * z;
    }
}

static double dist_nd(const double *a, const double *b, int dim) {
    double acc = 0.0;
    for (int k = 0; k < dim; ++k) {
        const double d = a[k] - b[k];
        acc = k == 0 ? d * d : acc + d * d;
    }
    return sqrt(acc);
}

/* Source proposal of the Green-source mixture: the source's Gaussian
 * exp(-|x - c|^2 / R^2) as a normal of per-axis width sigma = R / sqrt 2,
 * truncated to proposal_box (all four 0: no truncation, i.e. +-inf limits);
 * its normalisation is the product over both axes of
 * sigma sqrt(pi/2) (erf((hi - c) / (sqrt2 sigma)) - erf((lo - c) / (sqrt2 sigma)))
 * (apps/welding/include/welding/gaussian_heat_source.hpp:116-126). */
typedef struct orc_proposal { double sigma, norm, lo[2], hi[2]; } orc_proposal;

static void mis_proposal(const wos_eq_desc *eq, orc_proposal *q) {
    const double sqrt_two = 1.41421356237309504880168872420969808;
    const double sqrt_pi_over_two = 1.25331413731550025120788264240552263;
    const double *b = eq->proposal_box;
    const int boxed = !(b[0] == 0.0 && b[1] == 0.0 && b[2] == 0.0 && b[3] == 0.0);
    q->sigma = eq->src_radius / sqrt(2.0);
    q->norm = 1.0;
    for (int k = 0; k < 2; ++k) {
        q->lo[k] = boxed ? b[2 * k] : -INFINITY;
        q->hi[k] = boxed ? b[2 * k + 1] : INFINITY;
        const double c = eq->src_center[k];
        q->norm *= q->sigma * sqrt_pi_over_two *
                   (erf((q->hi[k] - c) / (sqrt_two * q->sigma)) -
                    erf((q->lo[k] - c) / (sqrt_two * q->sigma)));
    }
}

static int in_box(const orc_proposal *q, const double *x) {
    return x[0] >= q->lo[0] && x[0] <= q->hi[0] && x[1] >= q->lo[1] && x[1] <= q->hi[1];
}

/* Green-source mixture estimate on one ball, balance heuristic
 * (include/wos/source_sampling.hpp:35-78) with the finite-box Gaussian
 * proposal sampled by Box-Muller rejection and its density
 * (gaussian_heat_source.hpp:45-74, :109-113).  NaN where the reference
 * throws (rejection exhausted, mixture density not positive). */
static double mis_term(const wos_eq_desc *eq, const orc_proposal *q, const double *c, double R,
                       orc_rng *rng) {
    const double beta = eq->mis_green_probability;
    double sp[2];
    if (rng_unit(rng) < beta) {
        eq_sample_green(eq, 2, c, R, rng, sp);
    } else {
        int ok = 0;
        for (int attempt = 0; attempt < 1000000; ++attempt) {
            const double magnitude = sqrt(-2.0 * log(rng_unit_open(rng)));
            const double angle = 2.0 * ORC_PI * rng_unit(rng);
            sp[0] = eq->src_center[0] + q->sigma * magnitude * cos(angle);
            sp[1] = eq->src_center[1] + q->sigma * magnitude * sin(angle);
            if (in_box(q, sp)) { ok = 1; break; }
        }
        if (!ok) return NAN;
    }
    const double d = dist_nd(c, sp, 2);
    if (!(d > 0.0 && d < R)) return 0.0;
    const double mass = eq_green_mass(eq, 2, R);
    const double G = eq_green(eq, 2, R, d);
    const double green_pdf = G / mass;
    double source_pdf = 0.0;
    if (in_box(q, sp)) {
        const double ux = (sp[0] - eq->src_center[0]) / q->sigma;
        const double uy = (sp[1] - eq->src_center[1]) / q->sigma;
        source_pdf = exp(-0.5 * (ux * ux + uy * uy)) / q->norm;
    }
    const double mixture = beta * green_pdf + (1.0 - beta) * source_pdf;
    if (!(mixture > 0.0) || !isfinite(mixture)) return NAN;
    const double f = eq_source(eq, sp, 2);
    return f == 0.0 ? 0.0 : G * f / mixture;
}

typedef struct orc_sample { double value; int steps; int hit; } orc_sample;

/* Opt-in tally of the algorithmic work of the walk-step queries (node
 * records read, primitive tests, queries), summed over all threads. */
static int g_count_on = 0;
static int64_t g_count_nodes = 0, g_count_prims = 0, g_count_queries = 0;

void orc_count_reset(int on) {
    g_count_on = on;
    g_count_nodes = g_count_prims = g_count_queries = 0;
}
void orc_count_get(int64_t out[3]) {
    out[0] = g_count_nodes; out[1] = g_count_prims; out[2] = g_count_queries;
}

/* include/wos/wos.hpp:54-138 */
static orc_sample wos_sample(const orc_bvh *b, const wos_eq_desc *eq, const double *p0,
                             const orc_npq *start, double epsilon, int max_steps,
                             orc_rng *walk, orc_rng *source) {
    const int dim = b->dim;
    double p[3] = {p0[0], p0[1], dim == 3 ? p0[2] : 0.0};
    orc_npq q = *start;
    double R = start->distance;
    double sample = 0.0, weight = 1.0;
    int steps = 0;
    const int has_source = eq->source_kind != WOS_SRC_NONE;
    const int mis = has_source && dim == 2 && eq->source_mode == WOS_SOURCE_GREEN_MIS;
    orc_proposal prop = {0.0, 0.0, {0.0, 0.0}, {0.0, 0.0}};
    if (mis) mis_proposal(eq, &prop);
    while (R > epsilon && steps < max_steps) {
        rng_begin_step(walk, (uint32_t)steps, 0);
        rng_begin_step(source, (uint32_t)steps, (uint32_t)(dim - 1));
        if (has_source) {
            const int may = dim == 2 ? eq_source_may_intersect(eq, p, R) : 1;
            if (may) {
                double sp[3] = {0, 0, 0};
                if (mis) {
                    sample += weight * mis_term(eq, &prop, p, R, source);
                } else if (eq->source_mode == WOS_SOURCE_GREEN) {
                    eq_sample_green(eq, dim, p, R, source, sp);
                    const double f = eq_source(eq, sp, dim);
                    if (f != 0.0) sample += weight * eq_green_mass(eq, dim, R) * f;
                } else {
                    sphere_sample(dim, p, R, source, sp);
                    const double f = eq_source(eq, sp, dim);
                    if (f != 0.0)
                        sample += weight * sphere_volume(dim, R) * f *
                                  eq_green(eq, dim, R, dist_nd(p, sp, dim));
                }
            }
        }
        if (eq->kind == WOS_EQ_SCREENED)
            weight *= dim == 2 ? orc_screened_weight_2d(eq->alpha, R)
                               : orc_screened_weight_3d(eq->alpha, R);
        else if (eq->kind == WOS_EQ_HELMHOLTZ)
            weight *= dim == 2 ? orc_helmholtz_weight_2d(eq->alpha, R)
                               : orc_helmholtz_weight_3d(eq->alpha, R);
        double np[3] = {0, 0, 0};
        step_on_sphere(dim, p, R, walk, np);
        memcpy(p, np, sizeof(p));
        ++steps;
        if (g_count_on) {
            int64_t nr = 0, pt = 0;
            orc_bvh_npq(b, p, &q, &nr, &pt);
            __atomic_fetch_add(&g_count_nodes, nr, __ATOMIC_RELAXED);
            __atomic_fetch_add(&g_count_prims, pt, __ATOMIC_RELAXED);
            __atomic_fetch_add(&g_count_queries, 1, __ATOMIC_RELAXED);
        } else {
            orc_bvh_npq(b, p, &q, NULL, NULL);
        }
        R = q.distance;
    }
    orc_sample res;
    res.hit = R > epsilon;
    res.value = sample + weight * eq_boundary(eq, b, &q);
    res.steps = steps;
    return res;
}

/* include/wos/wos.hpp:160-191 (sequential Welford) with the per-point
 * streams of runner.hpp:208-221 (reference mode: 2D separate walk/source
 * xoshiro streams, 3D one shared stream) or the counter stream. */
void orc_wos_run(const orc_bvh *b, const wos_eq_desc *eq, const double *p0, int n_walks,
                 double epsilon, int max_steps, int rng_mode, uint64_t point_seed,
                 double stats[4], int *max_steps_hits, double *walk_values, int *walk_steps) {
    orc_npq start;
    orc_bvh_npq(b, p0, &start, NULL, NULL);
    uint64_t xw[4], xs[4];
    orc_rng walk, source;
    memset(&walk, 0, sizeof walk);
    memset(&source, 0, sizeof source);
    walk.mode = source.mode = rng_mode;
    const uint64_t walk_seed = orc_splitmix64(point_seed ^ WOS_STREAM);
    if (rng_mode == ORC_RNG_REFERENCE) {
        orc_xoshiro_seed(xw, walk_seed);
        walk.xo = xw;
        if (b->dim == 2) {
            orc_xoshiro_seed(xs, orc_splitmix64(point_seed ^ SOURCE_STREAM));
            source.xo = xs;
        } else {
            source.xo = xw;
        }
    } else {
        walk.key[0] = source.key[0] = (uint32_t)walk_seed;
        walk.key[1] = source.key[1] = (uint32_t)(walk_seed >> 32);
    }
    double mean = 0.0, m2 = 0.0, total_steps = 0.0;
    int hits = 0;
    for (int i = 0; i < n_walks; ++i) {
        walk.walk = source.walk = (uint32_t)i;
        const orc_sample s = wos_sample(b, eq, p0, &start, epsilon, max_steps, &walk, &source);
        const double count = (double)(i + 1);
        const double delta_before = s.value - mean;
        mean += delta_before / count;
        const double delta_after = s.value - mean;
        m2 += delta_before * delta_after;
        total_steps += s.steps;
        if (s.hit) ++hits;
        if (walk_values) walk_values[i] = s.value;
        if (walk_steps) walk_steps[i] = s.steps;
    }
    const double var = n_walks > 1 ? m2 / (double)(n_walks - 1) : 0.0;
    stats[0] = mean;
    stats[1] = var;
    stats[2] = n_walks > 0 ? sqrt(var / (double)n_walks) : 0.0;
    stats[3] = n_walks > 0 ? total_steps / (double)n_walks : 0.0;
    *max_steps_hits = hits;
}

/* ===================================================================== */
/* sweeps                                                                 */
/* ===================================================================== */

typedef struct sweep_job {
    const orc_bvh *b;
    const wos_eq_desc *eq;
    const double *points;
    const int64_t *seed_index;
    int64_t n;
    int n_walks, max_steps, max_ray_attempts, rng_mode;
    double epsilon, tol;
    uint64_t seed;
    double *mean, *var, *se, *mean_steps;
    uint8_t *location;
    int32_t *hits, *amb;
    int64_t next;
    int64_t total_steps;
    pthread_mutex_t lock;
} sweep_job;

/* one point of the sweep body, runner.hpp:163-235 */
static int64_t sweep_point(sweep_job *J, int64_t i) {
    const int dim = J->b->dim;
    const double *p0 = J->points + (size_t)i * dim;
    const uint64_t lin = J->seed_index ? (uint64_t)J->seed_index[i] : (uint64_t)i;
    const uint64_t ps = orc_point_seed(J->seed, lin);
    orc_npq start;
    orc_bvh_npq(J->b, p0, &start, NULL, NULL);
    int amb = 0;
    const int loc = orc_classify(J->b, p0, start.distance, J->tol, J->max_ray_attempts,
                                 orc_splitmix64(ps ^ INSIDE_STREAM), &amb);
    double st[4] = {NAN, NAN, NAN, NAN};
    int hits = 0;
    int64_t steps = 0;
    if (loc == 2) {
        st[0] = eq_boundary(J->eq, J->b, &start);
        st[1] = st[2] = st[3] = 0.0;
    } else if (loc == 1) {
        orc_wos_run(J->b, J->eq, p0, J->n_walks, J->epsilon, J->max_steps, J->rng_mode, ps, st,
                    &hits, NULL, NULL);
        steps = (int64_t)llround(st[3] * (double)J->n_walks);
    }
    J->mean[i] = st[0];
    if (J->var) J->var[i] = st[1];
    if (J->se) J->se[i] = st[2];
    if (J->mean_steps) J->mean_steps[i] = st[3];
    if (J->location) J->location[i] = (uint8_t)loc;
    if (J->hits) J->hits[i] = hits;
    if (J->amb) J->amb[i] = amb;
    return steps;
}

static void *sweep_worker(void *arg) {
    sweep_job *J = (sweep_job *)arg;
    int64_t local_steps = 0;
    for (;;) {
        int64_t i = __atomic_fetch_add(&J->next, 1, __ATOMIC_RELAXED);
        if (i >= J->n) break;
        local_steps += sweep_point(J, i);
    }
    __atomic_fetch_add(&J->total_steps, local_steps, __ATOMIC_RELAXED);
    return NULL;
}

static void run_sweep(sweep_job *J, int n_threads) {
    if (n_threads <= 1) {
        for (int64_t i = 0; i < J->n; ++i) J->total_steps += sweep_point(J, i);
        return;
    }
    pthread_t th[256];
    if (n_threads > 256) n_threads = 256;
    for (int t = 0; t < n_threads; ++t) pthread_create(&th[t], NULL, sweep_worker, J);
    for (int t = 0; t < n_threads; ++t) pthread_join(th[t], NULL);
}

int orc_solve_points(const orc_bvh *b, const wos_eq_desc *eq, const double *points, int64_t n,
                     const int64_t *seed_index, int n_walks, double epsilon, int max_steps,
                     int max_ray_attempts, uint64_t seed, int rng_mode, int n_threads,
                     double *mean, double *var, double *se, double *mean_steps,
                     uint8_t *location, int32_t *hits, int32_t *amb_retries,
                     int64_t *total_steps) {
    double g[6];
    orc_mesh_bbox(b, g);
    double scale = 1.0;
    for (int k = 0; k < 3; ++k) if (g[2 * k + 1] - g[2 * k] > scale) scale = g[2 * k + 1] - g[2 * k];
    sweep_job J;
    memset(&J, 0, sizeof J);
    J.b = b; J.eq = eq; J.points = points; J.seed_index = seed_index; J.n = n;
    J.n_walks = n_walks; J.max_steps = max_steps; J.max_ray_attempts = max_ray_attempts;
    J.rng_mode = rng_mode; J.epsilon = epsilon; J.tol = 1e-10 * scale; J.seed = seed;
    J.mean = mean; J.var = var; J.se = se; J.mean_steps = mean_steps; J.location = location;
    J.hits = hits; J.amb = amb_retries;
    run_sweep(&J, n_threads);
    if (total_steps) *total_steps = J.total_steps;
    return 0;
}

/* src/include/wos/grid.hpp:12-18 */
static double grid_coordinate(double lo, double hi, int index, int count) {
    return count > 1 ? lo + (hi - lo) * (double)index / (double)(count - 1) : lo;
}

/* runner.hpp:82-238 on one worker set: grid over the mesh bbox,
 * tol = 1e-10 max(1, extent), point (i,j,k) seeded by its linear index */
int orc_solve_grid(const orc_bvh *b, const wos_eq_desc *eq, int nx, int ny, int nz,
                   int n_walks, double epsilon, int max_steps, int max_ray_attempts,
                   uint64_t seed, int rng_mode, int n_threads, double *mean, double *var,
                   double *se, double *mean_steps, uint8_t *location, int32_t *hits,
                   int32_t *amb_retries, int64_t *total_steps, double grid_bounds[6],
                   double *tol_out) {
    const int dim = b->dim;
    if (dim == 2) nz = 1;
    double g[6];
    orc_mesh_bbox(b, g);
    double scale = 1.0;
    for (int k = 0; k < 3; ++k) if (g[2 * k + 1] - g[2 * k] > scale) scale = g[2 * k + 1] - g[2 * k];
    const int64_t n = (int64_t)nx * ny * nz;
    double *pts = (double *)malloc(sizeof(double) * (size_t)n * dim);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const int64_t lin = ((int64_t)i * ny + j) * nz + k;
                pts[lin * dim + 0] = grid_coordinate(g[0], g[1], i, nx);
                pts[lin * dim + 1] = grid_coordinate(g[2], g[3], j, ny);
                if (dim == 3) pts[lin * dim + 2] = grid_coordinate(g[4], g[5], k, nz);
            }
    orc_solve_points(b, eq, pts, n, NULL, n_walks, epsilon, max_steps, max_ray_attempts, seed,
                     rng_mode, n_threads, mean, var, se, mean_steps, location, hits, amb_retries,
                     total_steps);
    free(pts);
    if (grid_bounds) memcpy(grid_bounds, g, sizeof g);
    if (tol_out) *tol_out = 1e-10 * scale;
    return 0;
}
