/*
 * wos_hip.h -- C ABI of the MI355X walk-on-spheres engine (libwos_hip.so).
 *
 * Plain C types only: host pointers, sizes and POD descriptors.  Every entry
 * point returns an int status (WOS_OK on success) and never throws; the text
 * of the last failure on the calling thread is available from
 * wos_last_error().  A context owns one GPU's copies of the mesh and BVH and
 * its scratch buffers; calls on one context are serialised by the caller.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   - the per-point sweep   run<N,Eq>            include/wos/runner.hpp:22-308
 *   - the estimator         wos_run/wos_sample   include/wos/wos.hpp:54-191
 *   - the classifier        classify_point       src/inside.cpp:196-260
 *   - the closest point     bvh_npq              src/bvh.cpp:295-387
 *   - the BVH build         build_bvh            src/bvh.cpp:234-253
 *   - the equation functors src/equations/{laplace,poisson,screened_poisson}.cpp
 *   - the plugin table      struct Equation      src/equations/equations.hpp:7-31
 */
#ifndef WOS_HIP_H
#define WOS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (reference: int 0 ok / 1 error, runner.hpp:29-64) ---- */
enum wos_status {
    WOS_OK = 0,
    WOS_ERR_INPUT = 1,   /* invalid argument or size                        */
    WOS_ERR_MESH = 2,    /* mesh missing, malformed or degenerate            */
    WOS_ERR_DEVICE = 3,  /* HIP runtime failure                              */
    WOS_ERR_IO = 4,      /* file create / write failure                      */
    WOS_ERR_STATE = 5    /* call order (e.g. solve before a mesh was set)    */
};

/* ---- equation descriptor ------------------------------------------------
 * A POD restatement of the reference's compile-time Eq functors
 * (src/equations/laplace.cpp:12-43, poisson.cpp:15-118,
 * screened_poisson.cpp:40-141) plus the data the benchmark problems need. */
enum wos_equation_kind {
    WOS_EQ_LAPLACE = 0,          /* Delta u = 0                                  */
    WOS_EQ_POISSON = 1,          /* Delta u = -f                                 */
    WOS_EQ_SCREENED = 2,         /* Delta u - alpha^2 u = -f                     */
    WOS_EQ_HELMHOLTZ = 3         /* Delta u + k^2 u = 0, k = alpha, no source: weights
                                    1/J0(kR) in 2D, kR/sin(kR) in 3D (helmholtz.cpp:15-78) */
};

/* values equal the reference SourceMode enum (include/wos/source_mode.hpp:5-9) */
enum wos_source_mode {
    WOS_SOURCE_UNIFORM = 0,      /* uniform ball point x Green function          */
    WOS_SOURCE_GREEN = 1,        /* Green-density radius, mass x f               */
    WOS_SOURCE_GREEN_MIS = 2     /* balance-heuristic mixture of the Green density and the
                                    source's own (Gaussian) density, chosen with probability
                                    mis_green_probability (source_sampling.hpp:35-78); 2D
                                    Gaussian sources of Poisson / screened Poisson only */
};

enum wos_source_kind {
    WOS_SRC_NONE = 0,            /* no source term, no draws (Laplace, screened 3D) */
    WOS_SRC_DISK_SQ = 1,         /* f = value if |x-c|^2 <= R^2 (poisson.cpp:27-41); 2D prefilter */
    WOS_SRC_DISK_SQRT = 2,       /* f = value if sqrt(|x-c|^2) <= R (screened_poisson.cpp:48-61); 2D prefilter */
    WOS_SRC_CONSTANT = 3,        /* f = value everywhere                          */
    WOS_SRC_GAUSSIAN = 4,        /* f = value * exp(-|x-c|^2 / R^2)               */
    WOS_SRC_ZERO = 5,            /* f = 0 but sampled (reference Poisson3D, poisson.cpp:79-82) */
    WOS_SRC_FIELD = 6            /* 2D: f = history + heat from *field_source (below), the
                                    sources of the welding levels 0-3 */
};

/* ---- field source (WOS_SRC_FIELD): the implicit-Euler heat sources of the
 * welding app (level0_cooling.cpp:194-196, level1_stationary_heat.cpp:146-148,
 * level2_transient_heat.cpp:255-264, level3_moving_heat.cpp:189-194):
 *   f(x) = history(x) + heat(x)
 *   history(x) = history_scale * v(x), v clamped below at 0 when
 *                clamp_nonnegative, v = FieldSampler2D interpolation of the
 *                previous field (field_sampler.hpp:51-117); no values: 0
 *   heat(x)    = heat_scale * (0 + sum_c a_c exp(-(dx^2 + dy^2) / 2) / b_c) / heat_divisor,
 *                dx = (x - cx_c) / sx_c, dy = (y - cy_c) / sy_c
 *                (gaussian_heat_source.hpp:38-41, moving_gaussian_heat_source.hpp:75-88);
 *                0 when heat_scale == 0 or there are no terms */
enum wos_interpolation {
    WOS_INTERP_BILINEAR = 0,     /* FieldInterpolation2D::Bilinear */
    WOS_INTERP_CUBIC = 1         /* tensor cubic Lagrange on 4 x 4 nodes (needs nx, ny >= 4,
                                    else bilinear, as the reference) */
};

#define WOS_MAX_HEAT_TERMS 4

typedef struct wos_heat_term {
    double cx, cy;               /* centre                                         */
    double sx, sy;               /* per-axis width sigma                           */
    double a, b;                 /* term = a * shape / b                           */
} wos_heat_term;

typedef struct wos_field_source {
    const double *values;        /* previous field, nx * ny, index i * ny + j (host); may be NULL */
    int32_t nx, ny;              /* grid nodes, >= 2 each                          */
    double bounds[4];            /* xmin, xmax, ymin, ymax of the grid             */
    int32_t interpolation;       /* enum wos_interpolation                         */
    int32_t clamp_nonnegative;   /* 1: max(0, v) (levels 2, 3); 0: v (level 0)     */
    double history_scale;        /* alpha^2 of the implicit-Euler step             */
    double heat_scale;           /* power fraction of the step (level 2), else 1   */
    double heat_divisor;         /* conductivity k                                 */
    int32_t n_heat_terms;        /* 0 .. WOS_MAX_HEAT_TERMS                        */
    int32_t reserved0;
    wos_heat_term heat[WOS_MAX_HEAT_TERMS];
} wos_field_source;

enum wos_boundary_kind {
    WOS_BND_X2Y2_ID0 = 0,        /* 2D: x^2-y^2 on boundary id 0, else 0 (laplace.cpp:16-18) */
    WOS_BND_Y5 = 1,              /* 3D: y5^0 + 4 y5^-3 harmonic (laplace.cpp:30-36)          */
    WOS_BND_ZERO = 2,            /* g = 0 (poisson.cpp:43-48)                                */
    WOS_BND_ID_CONST = 3,        /* g = bnd_values[boundary id] (2D)                         */
    WOS_BND_VERTEX = 4,          /* g = barycentric / linear blend of per-vertex values      */
    WOS_BND_SIN3X_COS2Y = 5,     /* g = sin(3x) cos(2y)                                      */
    WOS_BND_ONE = 6              /* g = 1 (helmholtz.cpp:22-26, :55-58)                      */
};

#define WOS_MAX_BND_VALUES 16

typedef struct wos_eq_desc {
    int32_t kind;                /* enum wos_equation_kind                     */
    int32_t source_mode;         /* enum wos_source_mode                       */
    int32_t source_kind;         /* enum wos_source_kind                       */
    int32_t boundary_kind;       /* enum wos_boundary_kind                     */
    double alpha;                /* screening coefficient (WOS_EQ_SCREENED)    */
    double src_value;            /* disk value / constant / Gaussian amplitude */
    double src_center[3];        /* disk centre / Gaussian centre              */
    double src_radius;           /* disk radius / Gaussian width sigma         */
    double bnd_values[WOS_MAX_BND_VALUES];  /* WOS_BND_ID_CONST               */
    int32_t n_bnd_values;
    int32_t reserved0;
    const double *vertex_values; /* WOS_BND_VERTEX: one value per mesh vertex (host) */
    int64_t n_vertex_values;
    /* WOS_SOURCE_GREEN_MIS: the source proposal is the normalised Gaussian of the
     * source (per-axis width src_radius / sqrt 2 about src_center), truncated to the
     * box when proposal_box = {xmin, xmax, ymin, ymax} is non-empty (all four 0: the
     * whole plane), as GaussianHeatSource2D (apps/welding/include/welding/
     * gaussian_heat_source.hpp:45-74); beta in (0, 1) is the Green-draw probability
     * (source_mis_green_probability, level1_stationary_heat.cpp:188-190) */
    double mis_green_probability;
    double proposal_box[4];
    const wos_field_source *field_source;   /* WOS_SRC_FIELD: the source (host memory) */
} wos_eq_desc;

/* ---- execution parameters ---------------------------------------------- */
enum wos_precision {
    WOS_FP64_PLAIN = 0,          /* the reference's FP64 arithmetic and search on counter
                                    streams (persistent lanes, one walk at a time)  */
    WOS_FP32_PERSISTENT = 1,     /* persistent lane-refilling walker, FP32 geometry, FP64 stats */
    WOS_FP64_REFERENCE = 2       /* FP64 on the reference's own per-point streams: a lane holds
                                    a point and runs its walks in order on xoshiro256++ seeded as in
                                    include/wos/runner.hpp:208-221 (2D: walk + source stream,
                                    3D: one shared stream), so every walk is the reference's
                                    walk (up to last-bit libm differences); no adjoint */
};

typedef struct wos_run_params {
    int32_t n_walks;             /* walks per inside point (reference --walks)     */
    int32_t max_steps;           /* reference --max-steps (wos.hpp:58)             */
    int32_t max_ray_attempts;    /* reference --max-ray-attempts (inside.cpp:228)  */
    int32_t precision;           /* enum wos_precision                             */
    double epsilon;              /* stopping distance (already screened-clamped)   */
    uint64_t seed;               /* global seed (runner.hpp:163-168)               */
    double target_se;            /* > 0: adaptive walks per point until SE <= target */
    int32_t max_walks;           /* adaptive cap per point                         */
    int32_t tune;                /* 0: the context's launch shape; 1 (FP32): the
                                    autotuned one (wos_ctx_autotune), found on
                                    first use of a problem; results are the same;
                                    2: also the tuned BVH width (the walks where
                                    the tree is searched may change, their
                                    expectation does not)                          */
    int32_t count_units;         /* 1: count the FP32 walker's query units into the
                                    summary (wavefront ballots; off by default)    */
    int32_t use_graph;           /* 1: run a fixed-walk FP32 grid solve as one
                                    captured hipGraph (classification, compaction,
                                    walker, Welford), captured on first use and
                                    replayed while problem, tuning and buffers stay
                                    the same; results equal use_graph = 0          */
} wos_run_params;

/* Per-point outputs, caller-allocated arrays of n entries (any may be NULL
 * except mean).  Matches the reference's per-point result policy
 * (runner.hpp:195-235): Outside/Indeterminate -> NaN, Boundary -> g, 0, 0, 0. */
typedef struct wos_point_stats {
    double *mean;
    double *variance;
    double *standard_error;
    double *mean_steps;
    uint8_t *location;           /* 0 outside, 1 inside, 2 boundary, 3 indeterminate */
    int32_t *max_steps_hits;
    int32_t *walks;              /* walks taken per point (adaptive mode)          */
    int32_t *ambiguous_ray_retries;
} wos_point_stats;

typedef struct wos_run_summary {
    int64_t inside_count;
    int64_t max_steps_hits;
    int64_t ambiguous_ray_retries;
    int64_t indeterminate_points;
    double sum_variance;         /* runner.hpp:226-228 sums over inside points     */
    double sum_standard_error;
    double sum_mean_steps;
    double wos_seconds;          /* device time of classify + walks + merge        */
    double walk_kernel_seconds;  /* device time of the walker kernel(s) alone      */
    int64_t total_walks;
    int64_t total_steps;
    double grid_bounds[6];       /* xmin,xmax,ymin,ymax,zmin,zmax (mesh bbox)      */
    double boundary_tolerance;   /* 1e-10 * max(1, extent) (runner.hpp:86-88)      */
    int32_t launch_blocks;
    int32_t launch_threads;
    int32_t lds_cached_nodes;    /* BVH nodes staged in LDS by the FP32 walker     */
    int32_t rounds;              /* walker launches (adaptive mode: > 1)           */
    /* FP32 query units, filled when count_units = 1 (else 0): */
    int64_t q_queries;           /* closest-point queries begun (walk steps and
                                    exact re-queries)                              */
    int64_t q_grid_bound;        /* ... answered by a distance-grid cell bound     */
    int64_t q_list_scans;        /* ... answered by a cell's primitive list        */
    int64_t q_list_entries;      /* cell-list entries read                         */
    int64_t q_list_prims;        /* primitives tested from cell lists              */
    int64_t q_nodes;             /* BVH internal nodes opened                      */
    int64_t q_leaf_prims;        /* primitives tested from BVH leaves              */
    int64_t failed_walks;        /* walks whose value is NaN: a rejection sampler
                                    gave up (screened radius samplers after 10^5
                                    trials, the GREEN_MIS proposal after 10^6; the
                                    reference loops until acceptance); their points'
                                    means are NaN                                   */
    double grid_build_seconds;   /* FP32: seconds spent building the distance grid
                                    in this call (0 when it was already built)      */
    int64_t grid_bytes;          /* FP32: device bytes of the distance grid         */
} wos_run_summary;

/* Launch shape of the FP32 persistent walker (set by the autotuner; a new
 * context starts from {128, -1, 64, 0, 40, 0.2, -1, 127, 8, 2, 0}). */
typedef struct wos_tuning {
    int32_t block_threads;       /* 64, 128, 192 or 256 walkers per block          */
    int32_t lds_nodes;           /* BVH nodes cached in LDS per block; -1 = all that
                                    fit without lowering occupancy                  */
    int32_t chunk;               /* walks a wavefront claims per refill            */
    int32_t blocks_per_cu;       /* persistent blocks per CU; 0 = occupancy limit  */
    int32_t min_ready;           /* lanes of a wavefront whose closest-point queries
                                    have finished before they take their walk step
                                    together (0..64; 64 = whole-wave lockstep)     */
    float relax;                 /* relative slack r of step queries, 0 <= r < 0.5:
                                    boxes farther than (1 - r) x the best distance
                                    are skipped and the sphere radius is the
                                    resulting lower bound (>= (1 - r) x distance);
                                    0 = exact search.  Walks are ended and projected
                                    with an exact search in every case.            */
    int32_t grid_cells;          /* cells per domain of the FP32 distance grid
                                    (cell lists near the boundary, cell distance
                                    bounds inside); 0 = no grid, -1 = sized per
                                    domain from its primitive count: 4096 cells
                                    per triangle or 512 per segment, rounded up to
                                    a power of two, at least 2^16 and at most 2^24
                                    (triangles) or 2^22 (segments)                  */
    int32_t grid_list_max;       /* longest cell list kept (1..127); a cell with
                                    more candidate primitives keeps none           */
    int32_t xcd_parts;           /* the walk items are split into this many
                                    contiguous parts (regions of the domain) and
                                    workgroup b starts on part b % xcd_parts, which
                                    keeps each XCD's L2 on one region; 8 = the
                                    XCDs of an MI355X, 1 = one shared range       */
    int32_t bvh_width;           /* 2 or 4: the FP32 walker's tree where the distance
                                    grid has no list (the binary tree, or the same
                                    tree with every second level folded into 4-wide
                                    nodes); the order of a relaxed search, and so a
                                    walk, can change with it, not its expectation  */
    int32_t stack_lds;           /* 0..64: entries of each lane's traversal stack
                                    kept in LDS, the rest in global memory; never
                                    changes a result                               */
} wos_tuning;

typedef struct wos_ctx wos_ctx;

/* Text of the last error reported on this thread ("" if none). */
const char *wos_last_error(void);

/* Library version string. */
const char *wos_version(void);

/* Create a context on HIP device `device`.  Returns NULL on failure. */
wos_ctx *wos_ctx_create(int device);
void wos_ctx_destroy(wos_ctx *ctx);

/* Set the domain: vertices (n_verts x dim doubles, row-major), primitives
 * (n_prims x dim vertex indices: segments in 2D, triangles in 3D) and, in
 * 2D, the boundary id of every segment (the reference's `l` record index,
 * src/mesh.cpp:156-212; NULL means all zero).  Builds the BVH on the host
 * with the reference split rule (src/bvh.cpp:174-253) and uploads FP64 and
 * FP32 device layouts.  Replaces load_mesh+build_bvh+bcast (runner.hpp:28-80). */
int wos_ctx_set_mesh(wos_ctx *ctx, int dim, const double *verts, int64_t n_verts,
                     const int32_t *prims, int64_t n_prims,
                     const int32_t *boundary_ids);

/* Several domains in one context (all of one dimension), for batched
 * solves: domain d is (verts[d], n_verts[d], prims[d], n_prims[d],
 * boundary_ids[d] or NULL).  Each gets its own reference-rule BVH; the
 * device arrays are concatenated (vertex indices of per-vertex data become
 * global, in domain order).  No reference counterpart: the reference runs one
 * mesh per process. */
int wos_ctx_set_domains(wos_ctx *ctx, int dim, int n_domains, const double *const *verts,
                        const int64_t *n_verts, const int32_t *const *prims,
                        const int64_t *n_prims, const int32_t *const *boundary_ids);

/* One Nx x Ny x Nz grid over every domain's own bounding box, all walked in
 * one launch (adaptive mode included).  Outputs are concatenated in domain
 * order (n_domains x Nx*Ny*Nz entries); point seeds restart per domain, as
 * in separate runs of the reference sweep.  domain_summaries: n_domains
 * entries or NULL; summary: totals or NULL. */
int wos_solve_grid_batch(wos_ctx *ctx, int nx, int ny, int nz, const wos_eq_desc *eq,
                         const wos_run_params *params, wos_point_stats *out,
                         wos_run_summary *domain_summaries, wos_run_summary *summary);

/* Host-only BVH build with the reference split rule, exported for
 * verification (replaces build_bvh, src/bvh.cpp:234-253; no device work).
 * Returns the node count (or -1 on error).  With non-NULL arrays it fills
 * per node: count (0 = internal), idx (leaf: first slot; internal: right
 * child), lo/hi (3 doubles each), and slots[n_prims] (leaf slot -> primitive). */
int64_t wos_bvh_build_host(int dim, const double *verts, int64_t n_verts,
                           const int32_t *prims, int64_t n_prims, int32_t *count,
                           int32_t *idx, double *lo, double *hi, int32_t *slots);

/* Closest boundary point of n points (n x dim doubles) of domain `domain`
 * (0 after wos_ctx_set_mesh): dist[i] = distance, nearest[i] (n x dim, may be
 * NULL) = the nearest point, prim[i] (may be NULL) = its primitive.
 * WOS_FP64_PLAIN runs the reference traversal and returns what bvh_npq
 * returns (include/wos/bvh.hpp:38-39, src/bvh.cpp).  WOS_FP32_PERSISTENT runs
 * the walker's FP32 query and projects onto the primitive it found in FP64;
 * with relax > 0 (FP32 only) boxes farther than (1 - relax) x the best
 * distance are skipped and lower[i] (may be NULL) receives the lower bound
 * of the distance that the walker would use as its sphere radius
 * (lower[i] = dist[i] for an exact search).  A relaxed search that the
 * distance grid answers from its cell bound alone reports an upper bound in
 * dist[i], prim[i] = -1 and a NaN nearest point. */
int wos_closest_points(wos_ctx *ctx, int domain, const double *points, int64_t n,
                       int precision, double relax, double *dist, double *lower,
                       double *nearest, int32_t *prim);

/* wos_closest_points without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results as wos_closest_points, bit for bit.  The
 * points and results are staged in one fixed 16 MiB device array per GPU
 * (held until the process ends), which limits n to 246 704 points in 3D and
 * 322 614 in 2D (larger n: WOS_ERR_INPUT).  The array is shared by every
 * call on that GPU, so calls that overlap on one GPU (two threads, two
 * contexts) race; serialise them.  FP32 needs the context's distance grid
 * already built (any earlier FP32 call builds it), else WOS_ERR_STATE. */
int wos_closest_points_noalloc(wos_ctx *ctx, int domain, const double *points, int64_t n,
                               int precision, double relax, double *dist, double *lower,
                               double *nearest, int32_t *prim);

/* Launch shape of the FP32 walker.  No reference counterpart (the reference
 * has one CPU code path); used by the autotuner. */
int wos_ctx_set_tuning(wos_ctx *ctx, const wos_tuning *tuning);
int wos_ctx_get_tuning(const wos_ctx *ctx, wos_tuning *tuning);

/* Autotune the FP32 walker's launch shape on an Nx x Ny x Nz grid problem
 * over every domain of the context (the problem wos_solve_grid_batch would
 * solve): pilot solves of about 16 M walks, coordinate descent over walkers
 * per block, lanes per walk-step phase, refill chunk, LDS node cache,
 * xcd_parts and the LDS share of the traversal stack (stack_lds).
 * These fields never change a result, so a tuned solve is bit-identical to
 * an untuned one; grid_cells, grid_list_max and relax are kept as set.  With
 * params->tune = 2 the search also takes the BVH width (2 or 4), which can
 * change walks where the tree is searched (not their expectation).  The
 * chosen shape becomes the context's tuning, is written to `chosen` (may be
 * NULL) and is cached for solves with params->tune = 1 or 2, which run this
 * search themselves on first use of a problem.  No reference counterpart. */
int wos_ctx_autotune(wos_ctx *ctx, int nx, int ny, int nz, const wos_eq_desc *eq,
                     const wos_run_params *params, wos_tuning *chosen);

/* Sweep an Nx x Ny x Nz grid over the mesh bounding box (Nz forced to 1 in
 * 2D), classify every point, walk every inside point and write row-major
 * (i, j, k) fields.  Replaces the body of run<N,Eq> (runner.hpp:82-238). */
int wos_solve_grid(wos_ctx *ctx, int nx, int ny, int nz, const wos_eq_desc *eq,
                   const wos_run_params *params, wos_point_stats *out,
                   wos_run_summary *summary);

/* Same estimator on an explicit point list (n x dim doubles).  Point i uses
 * the seed splitmix64(seed ^ splitmix64(i)).  n must be in [1, 2^31 - 1]
 * (WOS_ERR_INPUT otherwise: the reference's grids have at least one point,
 * src/main.cpp:233).  Replaces the direct wos::wos_run calls of the welding
 * apps (level1_stationary_heat.cpp:257). */
int wos_solve_points(wos_ctx *ctx, const double *points, int64_t n,
                     const wos_eq_desc *eq, const wos_run_params *params,
                     wos_point_stats *out, wos_run_summary *summary);

/* Point list with explicit seed indices: point i uses the seed
 * splitmix64(seed ^ splitmix64(seed_index[i])), the reference sweep's seed of
 * the grid point whose global linear index (i Ny + j) Nz + k is seed_index[i]
 * (include/wos/runner.hpp:163-168).  A subset of a grid solved this way gets
 * the grid solve's walks; this is how a rank of the reference's Cartesian
 * block decomposition (runner.hpp:103-134) seeds its block.  Otherwise as
 * wos_solve_points. */
int wos_solve_points_indexed(wos_ctx *ctx, const double *points, const int64_t *seed_index,
                             int64_t n, const wos_eq_desc *eq, const wos_run_params *params,
                             wos_point_stats *out, wos_run_summary *summary);

/* ---- adjoint (no reference counterpart: the reference has no gradients) --
 * Gradient of L with respect to the equation data, given dL/du for every
 * point of the solve (u = the per-point mean of the forward estimator).  The
 * walks are replayed from the counter streams (same seeds, same draws as the
 * forward solve with the same parameters) and differentiated pathwise:
 *   d value / d vertex data  = weight x blend coefficient at the exit
 *   d value / d bnd_values   = weight at an exit on that boundary id
 *   d value / d src_value, src_radius, src_center: derivative of each source term
 *   d value / d alpha (3D screened, uniform source sampling or no source):
 *       the weights' and Green functions' alpha-derivatives; NaN where the
 *       sampling depends on alpha (Green source mode, 2D).
 * grad_vertex: n_verts doubles (WOS_BND_VERTEX problems; zeros otherwise).
 * grad_params: WOS_N_GRAD doubles indexed by the WOS_GRAD_* offsets.
 * `out` receives the replayed forward statistics (out->mean required).
 * grad_vertex and the WOS_GRAD_BND_VALUES entries are sums over walks formed
 * with FP64 atomic adds, so their last bits can differ between runs (the
 * order of the adds); `out` and the other grad_params entries (per-point
 * sums added in point order) are bitwise reproducible. */
enum wos_grad_index {
    WOS_GRAD_ALPHA = 0,
    WOS_GRAD_SRC_VALUE = 1,
    WOS_GRAD_SRC_RADIUS = 2,
    WOS_GRAD_SRC_CENTER = 3,     /* 3 entries */
    WOS_GRAD_BND_VALUES = 6,     /* WOS_MAX_BND_VALUES entries */
    WOS_N_GRAD = 6 + WOS_MAX_BND_VALUES
};

int wos_adjoint_grid(wos_ctx *ctx, int nx, int ny, int nz, const wos_eq_desc *eq,
                     const wos_run_params *params, const double *dl_du, double *grad_vertex,
                     double *grad_params, wos_point_stats *out, wos_run_summary *summary);

int wos_adjoint_points(wos_ctx *ctx, const double *points, int64_t n, const wos_eq_desc *eq,
                       const wos_run_params *params, const double *dl_du, double *grad_vertex,
                       double *grad_params, wos_point_stats *out, wos_run_summary *summary);

#ifdef __cplusplus
}
#endif
#endif /* WOS_HIP_H */
