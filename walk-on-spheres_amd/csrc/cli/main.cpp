// main.cpp -- the `wos` command line tool on one MI355X.
//
// CLI: wos [options] [equation] [Nx Ny] [Nz] [mesh.obj]
// Options, defaults, messages and exit codes follow the reference launcher
// (src/main.cpp:22-393); one process drives one GPU instead of MPI ranks.
// Additive options select the GPU form and the benchmark problem data.
#include <cerrno>
#include <charconv>
#include <chrono>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <filesystem>
#include <system_error>
#include <vector>

#include "obj_mesh.hpp"
#include "registry.hpp"

using namespace woscli;

static void print_usage(const char *program) {
    std::printf(
        "Usage: %s [options] [equation] [Nx Ny] [Nz] [mesh.obj]\n"
        "\n"
        "Options:\n"
        "  -o, --output FILE  HDF5 output path (default: wos.h5)\n"
        "      --seed UINT64  fixed random seed (default: generated from current time)\n"
        "      --walks N      WoS paths per inside grid point (default: 10000)\n"
        "      --epsilon X    boundary stopping distance (default: 0.01)\n"
        "      --max-steps N  maximum steps per WoS path (default: 1000)\n"
        "      --max-ray-attempts N\n"
        "                     maximum 3D inside-test ray attempts (default: 12)\n"
        "      --alpha X      screened Poisson coefficient (default: 5;\n"
        "                     screened_poisson only)\n"
        "      --source-mode MODE\n"
        "                     Poisson source sampling: uniform or green\n"
        "                     (poisson and screened_poisson)\n"
        "      --force        allow an existing output file to be overwritten\n"
        "  -h, --help         show this help message\n"
        "\n"
        "GPU options:\n"
        "      --precision P  fp64 (one lane per walk, default), fp32 (persistent\n"
        "                     walker, FP32 geometry, FP64 statistics) or fp64-ref\n"
        "                     (FP64 on the reference's own per-point random streams)\n"
        "      --target-se X  add walks per point until its standard error <= X\n"
        "      --max-walks N  cap of walks per point with --target-se\n"
        "      --problem NAME replace the equation's built-in data: annulus,\n"
        "                     sphere-screened, star-gaussian, plate-gaussian\n"
        "      --device N     HIP device index (default: 0)\n",
        program);
}

static bool parse_int(const char *text, int minimum, int *value) {
    int parsed = 0;
    const char *end = text + std::strlen(text);
    const std::from_chars_result r = std::from_chars(text, end, parsed);
    if (r.ec != std::errc{} || r.ptr != end || parsed < minimum) return false;
    *value = parsed;
    return true;
}

static bool parse_uint64(const char *text, uint64_t *value) {
    uint64_t parsed = 0;
    const char *end = text + std::strlen(text);
    const std::from_chars_result r = std::from_chars(text, end, parsed);
    if (r.ec != std::errc{} || r.ptr != end) return false;
    *value = parsed;
    return true;
}

static bool parse_positive_double(const char *text, double *value) {
    if (*text == '\0' || std::isspace((unsigned char)*text)) return false;
    char *end = nullptr;
    errno = 0;
    const double parsed = std::strtod(text, &end);
    if (end == text || *end != '\0' || errno == ERANGE || !std::isfinite(parsed) || parsed <= 0.0)
        return false;
    *value = parsed;
    return true;
}

int main(int argc, char **argv) {
    const char *eq_name = "laplace";
    const char *mesh_filename = "meshes/annulus.obj";
    const char *output_filename = "wos.h5";
    int Nx = 32, Ny = 32, Nz = 32;
    int N_walks = 10000;
    double epsilon = 1e-2;
    int max_steps = 1000;
    int max_ray_attempts = 12;
    double alpha = 5.0;
    SourceMode source_mode = SourceMode::Uniform;
    bool force_output = false, seed_provided = false, alpha_provided = false,
         source_mode_provided = false;
    uint64_t seed = 0;
    RunOptions &opt = run_options();
    const char *problem_name = nullptr;

    std::vector<const char *> positional;
    auto need = [&](int &i, const char *what) -> bool {
        if (++i >= argc) {
            std::fprintf(stderr, "%s requires %s\n", argv[i - 1], what);
            return false;
        }
        return true;
    };
    for (int i = 1; i < argc; i++) {
        const char *a = argv[i];
        if (!std::strcmp(a, "-h") || !std::strcmp(a, "--help")) {
            print_usage(argv[0]);
            return 0;
        } else if (!std::strcmp(a, "-o") || !std::strcmp(a, "--output")) {
            if (!need(i, "a file path")) return 1;
            output_filename = argv[i];
        } else if (!std::strcmp(a, "--seed")) {
            if (!need(i, "an unsigned 64-bit integer")) return 1;
            if (!parse_uint64(argv[i], &seed)) {
                std::fprintf(stderr, "Invalid --seed value: %s\n", argv[i]);
                return 1;
            }
            seed_provided = true;
        } else if (!std::strcmp(a, "--walks")) {
            if (!need(i, "a positive integer")) return 1;
            if (!parse_int(argv[i], 1, &N_walks)) {
                std::fprintf(stderr, "Invalid --walks value: %s\n", argv[i]);
                return 1;
            }
        } else if (!std::strcmp(a, "--epsilon")) {
            if (!need(i, "a positive finite number")) return 1;
            if (!parse_positive_double(argv[i], &epsilon)) {
                std::fprintf(stderr, "Invalid --epsilon value: %s\n", argv[i]);
                return 1;
            }
        } else if (!std::strcmp(a, "--max-steps")) {
            if (!need(i, "a positive integer")) return 1;
            if (!parse_int(argv[i], 1, &max_steps)) {
                std::fprintf(stderr, "Invalid --max-steps value: %s\n", argv[i]);
                return 1;
            }
        } else if (!std::strcmp(a, "--max-ray-attempts")) {
            if (!need(i, "an integer of at least 3")) return 1;
            if (!parse_int(argv[i], 3, &max_ray_attempts)) {
                std::fprintf(stderr, "Invalid --max-ray-attempts value: %s\n", argv[i]);
                return 1;
            }
        } else if (!std::strcmp(a, "--alpha")) {
            if (!need(i, "a positive finite number")) return 1;
            if (!parse_positive_double(argv[i], &alpha)) {
                std::fprintf(stderr, "Invalid --alpha value: %s\n", argv[i]);
                return 1;
            }
            alpha_provided = true;
        } else if (!std::strcmp(a, "--source-mode")) {
            if (!need(i, "uniform or green")) return 1;
            if (!std::strcmp(argv[i], "uniform")) source_mode = SourceMode::Uniform;
            else if (!std::strcmp(argv[i], "green")) source_mode = SourceMode::Green;
            else {
                std::fprintf(stderr, "Invalid --source-mode value: %s\n", argv[i]);
                return 1;
            }
            source_mode_provided = true;
        } else if (!std::strcmp(a, "--force")) {
            force_output = true;
        } else if (!std::strcmp(a, "--precision")) {
            if (!need(i, "fp64, fp32 or fp64-ref")) return 1;
            if (!std::strcmp(argv[i], "fp64")) opt.precision = WOS_FP64_PLAIN;
            else if (!std::strcmp(argv[i], "fp32")) opt.precision = WOS_FP32_PERSISTENT;
            else if (!std::strcmp(argv[i], "fp64-ref")) opt.precision = WOS_FP64_REFERENCE;
            else {
                std::fprintf(stderr, "Invalid --precision value: %s\n", argv[i]);
                return 1;
            }
        } else if (!std::strcmp(a, "--target-se")) {
            if (!need(i, "a positive finite number")) return 1;
            if (!parse_positive_double(argv[i], &opt.target_se)) {
                std::fprintf(stderr, "Invalid --target-se value: %s\n", argv[i]);
                return 1;
            }
        } else if (!std::strcmp(a, "--max-walks")) {
            if (!need(i, "a positive integer")) return 1;
            if (!parse_int(argv[i], 1, &opt.max_walks)) {
                std::fprintf(stderr, "Invalid --max-walks value: %s\n", argv[i]);
                return 1;
            }
        } else if (!std::strcmp(a, "--problem")) {
            if (!need(i, "a problem name")) return 1;
            problem_name = argv[i];
        } else if (!std::strcmp(a, "--device")) {
            if (!need(i, "a device index")) return 1;
            if (!parse_int(argv[i], 0, &opt.device)) {
                std::fprintf(stderr, "Invalid --device value: %s\n", argv[i]);
                return 1;
            }
        } else if (a[0] == '-') {
            std::fprintf(stderr, "Unknown option: %s\n", a);
            return 1;
        } else {
            positional.push_back(a);
        }
    }
    if (positional.size() > 5) {
        std::fprintf(stderr, "Too many positional arguments\n");
        return 1;
    }
    if (positional.size() == 2) {
        std::fprintf(stderr, "Nx and Ny must be provided together\n");
        return 1;
    }
    if (!positional.empty()) eq_name = positional[0];
    if (positional.size() >= 3) {
        if (!parse_int(positional[1], 1, &Nx) || !parse_int(positional[2], 1, &Ny)) {
            std::fprintf(stderr, "Nx and Ny must be positive integers\n");
            return 1;
        }
    }
    if (positional.size() >= 4) {
        int parsed_nz = 0;
        if (parse_int(positional[3], 1, &parsed_nz)) {
            Nz = parsed_nz;
            if (positional.size() == 5) mesh_filename = positional[4];
        } else if (positional.size() == 5) {
            std::fprintf(stderr, "Nz must be a positive integer when a fifth positional argument is present\n");
            return 1;
        } else {
            mesh_filename = positional[3];
        }
    }
    if (!seed_provided)
        seed = (uint64_t)std::chrono::high_resolution_clock::now().time_since_epoch().count();
    std::printf("Random seed: %llu%s\n", (unsigned long long)seed,
                seed_provided ? " (fixed by --seed)" : " (generated from current time)");

    const Equation *eq = nullptr;
    for (const Equation *e : equation_registry)
        if (!std::strcmp(e->name, eq_name)) { eq = e; break; }
    if (!eq) {
        std::fprintf(stderr, "Unknown equation: '%s'\nAvailable equations:\n", eq_name);
        for (const Equation *e : equation_registry) std::fprintf(stderr, "  - %s\n", e->name);
        return 1;
    }
    if (alpha_provided && eq != &screened_poisson)
        std::fprintf(stderr, "Warning: --alpha applies only to screened_poisson; "
                             "it will be ignored for equation '%s'.\n", eq->name);
    if (source_mode_provided && eq != &poisson && eq != &screened_poisson)
        std::fprintf(stderr, "Warning: --source-mode applies only to poisson and "
                             "screened_poisson; it will be ignored for equation '%s'.\n", eq->name);
    if (problem_name) {
        int kind = -1;
        opt.problem = named_problem(problem_name, &kind);
        const int want = eq == &laplace ? WOS_EQ_LAPLACE : eq == &poisson ? WOS_EQ_POISSON : WOS_EQ_SCREENED;
        if (!opt.problem || kind != want) {
            std::fprintf(stderr, "Unknown problem '%s' for equation '%s'\n", problem_name, eq->name);
            return 1;
        }
    }

    if (!force_output) {
        std::error_code error;
        const bool exists = std::filesystem::exists(output_filename, error);
        if (error) {
            std::fprintf(stderr, "Could not inspect output path %s: %s\n", output_filename,
                         error.message().c_str());
            return 1;
        }
        if (exists) {
            std::fprintf(stderr, "Refusing to overwrite existing output: %s\n"
                                 "Choose another path with --output, or pass --force.\n",
                         output_filename);
            return 1;
        }
    }

    int dim = 0;
    try {
        dim = peek_mesh_dim(mesh_filename);
    } catch (const std::exception &error) {
        std::fprintf(stderr, "%s\n", error.what());
        return 1;
    }
    const int rank = 0, size = 1;
    return dim == 2 ? eq->run_2D(rank, size, mesh_filename, output_filename, Nx, Ny, Nz, N_walks,
                                 epsilon, max_steps, max_ray_attempts, seed, alpha, source_mode)
                    : eq->run_3D(rank, size, mesh_filename, output_filename, Nx, Ny, Nz, N_walks,
                                 epsilon, max_steps, max_ray_attempts, seed, alpha, source_mode);
}
