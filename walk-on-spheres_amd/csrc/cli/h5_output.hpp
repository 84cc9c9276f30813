// h5_output.hpp -- result file of the `wos` command line tool.
//
// Layout of the reference writer (include/wos/hdf5_io.hpp:147-322):
//   /mean /variance /standard_error /mean_steps   f64 [Nx][Ny][Nz]
//   /location                                     u8  [Nx][Ny][Nz]
//   /x /y /z                                      f64 grid coordinates
//   /metadata/N_walks (int) epsilon (f64) max_steps (int)
//             max_steps_hits (long long) seed (u64)
//   3D also:  max_ray_attempts (int) ambiguous_ray_retries (long long)
//             indeterminate_points (long long)
//   source metadata (poisson, screened_poisson): alpha (f64) source_mode (int)
// One process writes the whole file (serial HDF5) instead of MPI-IO blocks.
#pragma once
#include <cstdint>
#include <vector>

namespace woscli {

struct H5Result {
    int nx, ny, nz;
    double bounds[6];  // xmin, xmax, ymin, ymax, zmin, zmax
    const double *mean, *variance, *standard_error, *mean_steps;
    const uint8_t *location;
    int n_walks;
    double epsilon;
    int max_steps;
    long long max_steps_hits;
    bool ray_metadata;
    int max_ray_attempts;
    long long ambiguous_ray_retries;
    long long indeterminate_points;
    uint64_t seed;
    bool source_metadata;
    double alpha;
    int source_mode;
};

// Returns false after printing "HDF5 error while <operation>" to stderr.
bool write_h5(const char *path, const H5Result &r);

}  // namespace woscli
