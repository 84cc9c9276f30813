// registry.hpp -- equation registry of the `wos` command line tool.
//
// Same plugin table as the reference (src/equations/equations.hpp:7-31):
// name + run_2D + run_3D with the reference's parameter list, looked up by
// exact name.  The GPU-specific choices (precision, adaptive target, problem
// data, device) travel in RunOptions, set by main (or wos_hip_run) before
// dispatch on the thread that runs the equation; each thread has its own.
#pragma once
#include <cstdint>

#include "../../../include/wos_hip.h"

namespace woscli {

enum class SourceMode : int { Uniform = 0, Green = 1 };

struct Equation {
    const char *name;
    int (*run_2D)(int rank, int size, const char *mesh_filename, const char *output_filename,
                  int Nx, int Ny, int Nz, int N_walks, double epsilon, int max_steps,
                  int max_ray_attempts, uint64_t seed, double alpha, SourceMode source_mode);
    int (*run_3D)(int rank, int size, const char *mesh_filename, const char *output_filename,
                  int Nx, int Ny, int Nz, int N_walks, double epsilon, int max_steps,
                  int max_ray_attempts, uint64_t seed, double alpha, SourceMode source_mode);
};

struct RunOptions {
    int precision = WOS_FP64_PLAIN;
    double target_se = 0.0;
    int max_walks = 0;
    int device = 0;
    const wos_eq_desc *problem = nullptr;  // replaces the built-in equation data
};

RunOptions &run_options();

extern const Equation laplace;
extern const Equation poisson;
extern const Equation screened_poisson;
extern const Equation helmholtz;

inline const Equation *const equation_registry[] = {&laplace, &poisson, &screened_poisson,
                                                     &helmholtz};

// Descriptors of named benchmark problems (--problem); nullptr if unknown.
const wos_eq_desc *named_problem(const char *name, int *kind);

}  // namespace woscli
