/*
 * wos_oracle.h -- CPU restatement of the reference walk-on-spheres path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline).  The product (libwos_hip.so) never links it.
 *
 * Plain C11, FP64, compiled with -ffp-contract=off so that its arithmetic is
 * the reference's arithmetic operation for operation (the reference builds
 * with g++ -O3 for baseline x86-64, which has no FMA).  Every function cites
 * the reference file:line it restates (paths relative to the reference repo).
 *
 * Parity pins (see tests/): splitmix64/xoshiro256++/npq/eberly/winding/
 * moller_trumbore/classify/wos_run are checked bit-for-bit against the
 * compiled reference core (oracle/_ref) and against committed golden vectors;
 * philox4x32-10 against the published Random123 known-answer vectors;
 * i0/k0 against scipy.special.i0e/k0e; the estimator against closed-form
 * harmonic / screened solutions.
 */
#ifndef WOS_ORACLE_H
#define WOS_ORACLE_H

#include <stdint.h>
#include "../include/wos_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- random numbers ---------------- */
uint64_t orc_splitmix64(uint64_t x);
void orc_xoshiro_seed(uint64_t state[4], uint64_t seed);
uint64_t orc_xoshiro_next(uint64_t state[4]);
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
uint64_t orc_point_seed(uint64_t seed, uint64_t linear_index);

/* ---------------- special functions ---------------- */
double orc_i0_scaled(double x);
double orc_k0_scaled(double x);
double orc_log_i0(double x);
double orc_log_k0(double x);

/* Green functions (include/wos/poisson.hpp, include/wos/screened.hpp) */
double orc_poisson_green_2d(double R, double r);
double orc_poisson_green_3d(double R, double r);
double orc_screened_weight_2d(double alpha, double R);
double orc_screened_weight_3d(double alpha, double R);
double orc_screened_green_2d(double alpha, double R, double r);
double orc_screened_green_3d(double alpha, double R, double r);
double orc_screened_mass_2d(double alpha, double R);
double orc_screened_mass_3d(double alpha, double R);
/* Helmholtz (src/equations/helmholtz.cpp, include/wos/fastmath.hpp) */
double orc_bessel_j0_series(double x);
double orc_helmholtz_weight_2d(double k, double R);
double orc_helmholtz_weight_3d(double k, double R);

/* ---------------- primitives ---------------- */
double orc_npq_seg(const double s0[2], const double s1[2], const double p[2],
                   double nearest[2], double *par);
double orc_eberly(const double p[3], const double v0[3], const double v1[3],
                  const double v2[3], double nearest[3], double st[2]);
int orc_moller_trumbore(const double p[3], const double ray[3], const double a[3],
                        const double b[3], const double c[3]);

/* ---------------- mesh + BVH ---------------- */
typedef struct orc_node {
    double bmin[3], bmax[3];
    int prim_count;   /* 0: internal node                                  */
    int idx;          /* leaf: first slot in prims; internal: right child  */
} orc_node;

typedef struct orc_bvh {
    int dim, n_verts, n_prims, n_nodes;
    double *verts;    /* n_verts x dim */
    int *mesh_prims;  /* n_prims x dim */
    int *bids;        /* n_prims (2D) */
    orc_node *nodes;
    int *prims;       /* leaf slot -> primitive id */
} orc_bvh;

typedef struct orc_npq {
    double distance;
    double point[3];
    int prim;         /* nearest primitive id                             */
    int bid;          /* 2D boundary id of that primitive, -1 in 3D       */
    double s, t;      /* 2D: segment parameter in s; 3D: Eberly (s, t)    */
} orc_npq;

orc_bvh *orc_bvh_build(int dim, const double *verts, int n_verts,
                       const int *prims, int n_prims, const int *bids);
void orc_bvh_free(orc_bvh *bvh);
int orc_bvh_n_nodes(const orc_bvh *bvh);
void orc_bvh_export(const orc_bvh *bvh, int *prim_count, int *idx, double *bmin,
                    double *bmax, int *leaf_prims);
void orc_bvh_npq(const orc_bvh *bvh, const double *p, orc_npq *out,
                 int64_t *node_reads, int64_t *prim_tests);
void orc_npq_brute(const orc_bvh *bvh, const double *p, orc_npq *out);
void orc_mesh_bbox(const orc_bvh *bvh, double bounds[6]);

/* ---------------- classification ---------------- */
int orc_winding_number(const orc_bvh *bvh, const double p[2]);
int orc_ray_cast(const orc_bvh *bvh, const double p[3], const double ray[3],
                 int *ambiguous);
int orc_classify(const orc_bvh *bvh, const double *p, double distance, double tol,
                 int max_ray_attempts, uint64_t inside_seed, int *ambiguous_retries);

/* ---------------- estimator ---------------- */
enum { ORC_RNG_REFERENCE = 0, ORC_RNG_COUNTER = 1 };
/* The first n uniforms (FP64) of the counter stream of step `step` of walk
 * `walk` under key `stream_seed`, from slot `first` on. */
void orc_counter_draws(uint64_t stream_seed, uint32_t walk, uint32_t step, uint32_t first,
                       int n, double *out);

/* opt-in tally of walk-step query work: {node records read, primitive
 * tests, queries} */
void orc_count_reset(int on);
void orc_count_get(int64_t out[3]);

/* FieldSampler2D interpolation and the welding source f = history + heat */
double orc_field_sample(const wos_field_source *f, double x, double y);
double orc_field_source_value(const wos_field_source *f, double x, double y);

void orc_wos_run(const orc_bvh *bvh, const wos_eq_desc *eq, const double *p0,
                 int n_walks, double epsilon, int max_steps, int rng_mode,
                 uint64_t point_seed, double stats[4], int *max_steps_hits,
                 double *walk_values, int *walk_steps);

int orc_solve_points(const orc_bvh *bvh, const wos_eq_desc *eq,
                     const double *points, int64_t n, const int64_t *seed_index,
                     int n_walks, double epsilon, int max_steps,
                     int max_ray_attempts, uint64_t seed, int rng_mode,
                     int n_threads, double *mean, double *var, double *se,
                     double *mean_steps, uint8_t *location, int32_t *hits,
                     int32_t *amb_retries, int64_t *total_steps);

int orc_solve_grid(const orc_bvh *bvh, const wos_eq_desc *eq, int nx, int ny, int nz,
                   int n_walks, double epsilon, int max_steps, int max_ray_attempts,
                   uint64_t seed, int rng_mode, int n_threads, double *mean,
                   double *var, double *se, double *mean_steps, uint8_t *location,
                   int32_t *hits, int32_t *amb_retries, int64_t *total_steps,
                   double grid_bounds[6], double *tol);

#ifdef __cplusplus
}
#endif
#endif
