// run_api.cpp -- wos_hip_run: the registry's run_2D / run_3D behind one C
// symbol (libwos_run.so), for callers that dispatch like the reference's
// main (src/main.cpp:380-389).
#include <cstdio>

#include "../../../include/wos_run.h"
#include "registry.hpp"

namespace {
constexpr int kEquations = (int)(sizeof(woscli::equation_registry) /
                                 sizeof(woscli::equation_registry[0]));
}

extern "C" {

int wos_hip_equation_count(void) { return kEquations; }

const char *wos_hip_equation_name(int equation_id) {
    if (equation_id < 0 || equation_id >= kEquations) return nullptr;
    return woscli::equation_registry[equation_id]->name;
}

int wos_hip_run(int dim, const char *mesh_filename, const char *output_filename, int Nx, int Ny,
                int Nz, int N_walks, double epsilon, int max_steps, int max_ray_attempts,
                uint64_t seed, double alpha, int source_mode, int equation_id,
                const wos_eq_desc *problem, const wos_exec_opts *opts) {
    if (dim != 2 && dim != 3) {
        std::fprintf(stderr, "dim must be 2 or 3\n");
        return 1;
    }
    if (equation_id < 0 || equation_id >= kEquations) {
        std::fprintf(stderr, "Unknown equation index %d\n", equation_id);
        return 1;
    }
    if (!mesh_filename || !output_filename) {
        std::fprintf(stderr, "mesh and output file names are required\n");
        return 1;
    }
    if (source_mode != 0 && source_mode != 1) {
        std::fprintf(stderr, "source_mode must be 0 (uniform) or 1 (green)\n");
        return 1;
    }
    woscli::RunOptions &o = woscli::run_options();
    const woscli::RunOptions saved = o;
    o.problem = problem;
    o.precision = opts ? opts->precision : WOS_FP64_PLAIN;
    o.device = opts ? opts->device : 0;
    o.target_se = opts ? opts->target_se : 0.0;
    o.max_walks = opts ? opts->max_walks : 0;
    const woscli::Equation *e = woscli::equation_registry[equation_id];
    const auto mode = static_cast<woscli::SourceMode>(source_mode);
    const int rc = dim == 2 ? e->run_2D(0, 1, mesh_filename, output_filename, Nx, Ny, Nz,
                                        N_walks, epsilon, max_steps, max_ray_attempts, seed,
                                        alpha, mode)
                            : e->run_3D(0, 1, mesh_filename, output_filename, Nx, Ny, Nz,
                                        N_walks, epsilon, max_steps, max_ray_attempts, seed,
                                        alpha, mode);
    o = saved;
    return rc;
}

}  // extern "C"
