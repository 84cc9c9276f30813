// registry.cpp -- the equations of the `wos` tool and the sweep driver that
// runs them through libwos_hip.so.
//
// Equation data restates src/equations/laplace.cpp:12-43, poisson.cpp:15-118
// and screened_poisson.cpp:40-141; the driver prints the progress and summary
// lines of include/wos/runner.hpp:22-308 and writes the same HDF5 layout.
#include "registry.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <vector>

#include "h5_output.hpp"
#include "obj_mesh.hpp"

namespace woscli {

// one set of options per thread: wos_hip_run sets them on the calling
// thread for the duration of its call, so calls on several threads (each
// with its own context) do not see each other's precision or device
RunOptions &run_options() {
    thread_local RunOptions opts;
    return opts;
}

namespace {

wos_eq_desc blank(int kind) {
    wos_eq_desc e;
    std::memset(&e, 0, sizeof e);
    e.kind = kind;
    return e;
}

// run<N,Eq> of include/wos/runner.hpp with the GPU engine
int run(int dim, const char *mesh_file, const char *out_file, int Nx, int Ny, int Nz,
        int N_walks, double epsilon, int max_steps, int max_ray_attempts, uint64_t seed,
        const wos_eq_desc &eq, bool source_metadata, double alpha_md, int source_mode) {
    const RunOptions &opt = run_options();
    std::printf("Reading mesh from disk (%s)...\n", mesh_file);
    ObjMesh mesh;
    try {
        mesh = load_mesh(mesh_file, dim);
    } catch (const std::exception &err) {
        std::fprintf(stderr, "%s\n", err.what());
        return 1;
    }
    std::printf("Finished reading. Broadcasting mesh...\n");
    std::printf("Finished broadcasting.\n");
    if (eq.boundary_kind == WOS_BND_VERTEX && eq.n_vertex_values != mesh.n_verts()) {
        std::fprintf(stderr, "per-vertex boundary data needs %lld values\n",
                     (long long)mesh.n_verts());
        return 1;
    }

    std::printf("Building BVH...\n");
    wos_ctx *ctx = wos_ctx_create(opt.device);
    if (!ctx) {
        std::fprintf(stderr, "BVH build failed: %s\n", wos_last_error());
        return 1;
    }
    const auto t0 = std::chrono::steady_clock::now();
    int rc = wos_ctx_set_mesh(ctx, dim, mesh.verts.data(), mesh.n_verts(), mesh.prims.data(),
                              mesh.n_prims(), dim == 2 ? mesh.bids.data() : nullptr);
    const double build_s =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rc != WOS_OK) {
        std::fprintf(stderr, "BVH build failed: %s\n", wos_last_error());
        wos_ctx_destroy(ctx);
        return 1;
    }
    std::printf("Finished BVH build in %.6f s.\n", build_s);

    if (dim == 2) Nz = 1;
    std::printf("Parallelisation: %d x %d x %d ranks across (Nx=%d, Ny=%d, Nz=%d)\n", 1, 1, 1,
                Nx, Ny, Nz);
    const size_t n = (size_t)Nx * Ny * Nz;
    std::vector<double> mean(n), var(n), se(n), msteps(n);
    std::vector<uint8_t> loc(n);
    wos_point_stats st{mean.data(), var.data(), se.data(), msteps.data(), loc.data(),
                       nullptr, nullptr, nullptr};
    wos_run_params prm{};
    prm.n_walks = N_walks;
    prm.max_steps = max_steps;
    prm.max_ray_attempts = max_ray_attempts;
    prm.precision = opt.precision;
    prm.epsilon = epsilon;
    prm.seed = seed;
    prm.target_se = opt.target_se;
    prm.max_walks = opt.max_walks > 0 ? opt.max_walks : N_walks;
    wos_run_summary sum{};
    std::printf("Beginning walk-on-spheres...\n");
    const auto t1 = std::chrono::steady_clock::now();
    rc = wos_solve_grid(ctx, Nx, Ny, Nz, &eq, &prm, &st, &sum);
    const double wos_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
    wos_ctx_destroy(ctx);
    if (rc != WOS_OK) {
        std::fprintf(stderr, "walk-on-spheres failed: %s\n", wos_last_error());
        return 1;
    }
    if (sum.inside_count > 0) {
        const double count = (double)sum.inside_count;
        std::printf("\nWoS inside-grid averages:\n");
        std::printf("  number of inside points:       %lld\n", (long long)sum.inside_count);
        std::printf("  average variance:              %.12g\n", sum.sum_variance / count);
        std::printf("  average standard error:        %.12g\n", sum.sum_standard_error / count);
        std::printf("  average steps per walk:        %.12g\n", sum.sum_mean_steps / count);
        std::printf("  paths reaching max steps:      %lld\n", (long long)sum.max_steps_hits);
        std::printf("\n");
    }
    if (dim == 3) {
        std::printf("3D ray-classification diagnostics:\n");
        std::printf("  ambiguous ray retries:         %lld\n", (long long)sum.ambiguous_ray_retries);
        std::printf("  indeterminate grid points:     %lld\n\n", (long long)sum.indeterminate_points);
    }
    std::printf("Finished walk-on-spheres in %.6f s. Writing results...\n", wos_s);

    H5Result r{};
    r.nx = Nx; r.ny = Ny; r.nz = Nz;
    for (int k = 0; k < 6; ++k) r.bounds[k] = sum.grid_bounds[k];
    r.mean = mean.data(); r.variance = var.data(); r.standard_error = se.data();
    r.mean_steps = msteps.data(); r.location = loc.data();
    r.n_walks = N_walks; r.epsilon = epsilon; r.max_steps = max_steps;
    r.max_steps_hits = sum.max_steps_hits;
    r.ray_metadata = dim == 3;
    r.max_ray_attempts = max_ray_attempts;
    r.ambiguous_ray_retries = sum.ambiguous_ray_retries;
    r.indeterminate_points = sum.indeterminate_points;
    r.seed = seed;
    r.source_metadata = source_metadata;
    r.alpha = alpha_md;
    r.source_mode = source_mode;
    if (!write_h5(out_file, r)) {
        std::fprintf(stderr, "Failed writing results to %s.\n", out_file);
        return 1;
    }
    std::printf("Finished writing to %s.\n", out_file);
    return 0;
}

const wos_eq_desc *override_or(const wos_eq_desc &builtin) {
    const wos_eq_desc *p = run_options().problem;
    return p ? p : &builtin;
}

// ---- laplace (src/equations/laplace.cpp)
int laplace_run(int dim, const char *mesh, const char *out, int Nx, int Ny, int Nz, int N_walks,
                double eps, int max_steps, int mra, uint64_t seed) {
    wos_eq_desc e = blank(WOS_EQ_LAPLACE);
    e.boundary_kind = dim == 2 ? WOS_BND_X2Y2_ID0 : WOS_BND_Y5;
    return run(dim, mesh, out, Nx, Ny, Nz, N_walks, eps, max_steps, mra, seed, *override_or(e),
               false, 0.0, 0);
}
int laplace_2d(int, int, const char *mesh, const char *out, int Nx, int Ny, int Nz, int N_walks,
               double eps, int max_steps, int mra, uint64_t seed, double, SourceMode) {
    return laplace_run(2, mesh, out, Nx, Ny, Nz, N_walks, eps, max_steps, mra, seed);
}
int laplace_3d(int, int, const char *mesh, const char *out, int Nx, int Ny, int Nz, int N_walks,
               double eps, int max_steps, int mra, uint64_t seed, double, SourceMode) {
    return laplace_run(3, mesh, out, Nx, Ny, Nz, N_walks, eps, max_steps, mra, seed);
}

// ---- poisson (src/equations/poisson.cpp): metadata alpha is 0 (poisson.cpp:124-131)
int poisson_run(int dim, const char *mesh, const char *out, int Nx, int Ny, int Nz, int N_walks,
                double eps, int max_steps, int mra, uint64_t seed, SourceMode mode) {
    wos_eq_desc e = blank(WOS_EQ_POISSON);
    e.source_mode = (int)mode;
    if (dim == 2) {
        e.source_kind = WOS_SRC_DISK_SQ;
        e.src_value = 0.1;
        e.src_radius = 0.15;
        e.boundary_kind = WOS_BND_ZERO;
    } else {
        e.source_kind = WOS_SRC_ZERO;
        e.boundary_kind = WOS_BND_Y5;
    }
    wos_eq_desc use = *override_or(e);
    use.source_mode = (int)mode;
    return run(dim, mesh, out, Nx, Ny, Nz, N_walks, eps, max_steps, mra, seed, use, true, 0.0,
               (int)mode);
}
int poisson_2d(int, int, const char *mesh, const char *out, int Nx, int Ny, int Nz, int N_walks,
               double eps, int max_steps, int mra, uint64_t seed, double, SourceMode mode) {
    return poisson_run(2, mesh, out, Nx, Ny, Nz, N_walks, eps, max_steps, mra, seed, mode);
}
int poisson_3d(int, int, const char *mesh, const char *out, int Nx, int Ny, int Nz, int N_walks,
               double eps, int max_steps, int mra, uint64_t seed, double, SourceMode mode) {
    return poisson_run(3, mesh, out, Nx, Ny, Nz, N_walks, eps, max_steps, mra, seed, mode);
}

// ---- screened_poisson (src/equations/screened_poisson.cpp)
double screened_epsilon(double requested, double alpha) {
    const double limited = 0.05 / alpha;
    const double effective = std::min(requested, limited);
    if (effective < requested) {
        std::printf("Screened stopping distance reduced:\n"
                    "  requested epsilon:       %.12g\n"
                    "  effective epsilon:       %.12g\n"
                    "  alpha * effective eps:   %.12g\n",
                    requested, effective, alpha * effective);
    }
    return effective;
}
int screened_run(int dim, const char *mesh, const char *out, int Nx, int Ny, int Nz, int N_walks,
                 double eps, int max_steps, int mra, uint64_t seed, double alpha, SourceMode mode) {
    const double eff = screened_epsilon(eps, alpha);
    wos_eq_desc e = blank(WOS_EQ_SCREENED);
    e.alpha = alpha;
    e.source_mode = (int)mode;
    if (dim == 2) {
        e.source_kind = WOS_SRC_DISK_SQRT;
        e.src_value = 5.0;
        e.src_center[0] = 0.45;
        e.src_radius = 0.5;
        e.boundary_kind = WOS_BND_X2Y2_ID0;
    } else {
        e.source_kind = WOS_SRC_NONE;
        e.boundary_kind = WOS_BND_Y5;
    }
    wos_eq_desc use = *override_or(e);
    use.alpha = alpha;
    use.source_mode = (int)mode;
    return run(dim, mesh, out, Nx, Ny, Nz, N_walks, eff, max_steps, mra, seed, use, true, alpha,
               (int)mode);
}
int screened_2d(int, int, const char *mesh, const char *out, int Nx, int Ny, int Nz, int N_walks,
                double eps, int max_steps, int mra, uint64_t seed, double alpha, SourceMode mode) {
    return screened_run(2, mesh, out, Nx, Ny, Nz, N_walks, eps, max_steps, mra, seed, alpha, mode);
}
int screened_3d(int, int, const char *mesh, const char *out, int Nx, int Ny, int Nz, int N_walks,
                double eps, int max_steps, int mra, uint64_t seed, double alpha, SourceMode mode) {
    return screened_run(3, mesh, out, Nx, Ny, Nz, N_walks, eps, max_steps, mra, seed, alpha, mode);
}

// ---- helmholtz (src/equations/helmholtz.cpp): k = 1, no source, g = 1; the
// reference ignores --alpha and --source-mode here and writes no source
// metadata (helmholtz.cpp:79-94)
int helmholtz_run(int dim, const char *mesh, const char *out, int Nx, int Ny, int Nz,
                  int N_walks, double eps, int max_steps, int mra, uint64_t seed) {
    wos_eq_desc e = blank(WOS_EQ_HELMHOLTZ);
    e.alpha = 1.0;
    e.boundary_kind = WOS_BND_ONE;
    return run(dim, mesh, out, Nx, Ny, Nz, N_walks, eps, max_steps, mra, seed, *override_or(e),
               false, 0.0, 0);
}
int helmholtz_2d(int, int, const char *mesh, const char *out, int Nx, int Ny, int Nz, int N_walks,
                 double eps, int max_steps, int mra, uint64_t seed, double, SourceMode) {
    return helmholtz_run(2, mesh, out, Nx, Ny, Nz, N_walks, eps, max_steps, mra, seed);
}
int helmholtz_3d(int, int, const char *mesh, const char *out, int Nx, int Ny, int Nz, int N_walks,
                 double eps, int max_steps, int mra, uint64_t seed, double, SourceMode) {
    return helmholtz_run(3, mesh, out, Nx, Ny, Nz, N_walks, eps, max_steps, mra, seed);
}

}  // namespace

const Equation laplace = {"laplace", laplace_2d, laplace_3d};
const Equation poisson = {"poisson", poisson_2d, poisson_3d};
const Equation screened_poisson = {"screened_poisson", screened_2d, screened_3d};
const Equation helmholtz = {"helmholtz", helmholtz_2d, helmholtz_3d};

const wos_eq_desc *named_problem(const char *name, int *kind) {
    struct Named { wos_eq_desc annulus, sphere, star, plate; };
    // built once (thread-safe static initialisation)
    static const Named P = [] {
        Named n;
        n.annulus = blank(WOS_EQ_LAPLACE);         // g = 1 on boundary id 0, 0 on id 1
        n.annulus.boundary_kind = WOS_BND_ID_CONST;
        n.annulus.bnd_values[0] = 1.0;
        n.annulus.n_bnd_values = 2;
        n.sphere = blank(WOS_EQ_SCREENED);         // f = 1, g = 0
        n.sphere.source_kind = WOS_SRC_CONSTANT;
        n.sphere.src_value = 1.0;
        n.sphere.boundary_kind = WOS_BND_ZERO;
        n.star = blank(WOS_EQ_POISSON);            // f = exp(-|x|^2 / 0.1), g = sin 3x cos 2y
        n.star.source_kind = WOS_SRC_GAUSSIAN;
        n.star.src_value = 1.0;
        n.star.src_radius = std::sqrt(0.1);
        n.star.boundary_kind = WOS_BND_SIN3X_COS2Y;
        n.plate = blank(WOS_EQ_SCREENED);          // f = exp(-|x - x0|^2 / 9), g = 0
        n.plate.source_kind = WOS_SRC_GAUSSIAN;
        n.plate.src_value = 1.0;
        n.plate.src_radius = 3.0;
        n.plate.src_center[0] = 40.0;
        n.plate.src_center[1] = 30.0;
        n.plate.src_center[2] = 10.0;
        n.plate.boundary_kind = WOS_BND_ZERO;
        return n;
    }();
    const wos_eq_desc *p = nullptr;
    if (!std::strcmp(name, "annulus")) p = &P.annulus;
    else if (!std::strcmp(name, "sphere-screened")) p = &P.sphere;
    else if (!std::strcmp(name, "star-gaussian")) p = &P.star;
    else if (!std::strcmp(name, "plate-gaussian")) p = &P.plate;
    if (p && kind) *kind = p->kind;
    return p;
}

}  // namespace woscli
