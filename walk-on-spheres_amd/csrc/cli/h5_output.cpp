// h5_output.cpp -- serial HDF5 writer with the reference result layout.
#include "h5_output.hpp"

#include <hdf5.h>

#include <cstdio>

namespace woscli {

namespace {

double grid_coordinate(double lo, double hi, int i, int n) {
    return n > 1 ? lo + (hi - lo) * (double)i / (double)(n - 1) : lo;
}

bool write_field(hid_t file, const char *name, hid_t type, int nx, int ny, int nz, const void *data) {
    hsize_t dims[3] = {(hsize_t)nx, (hsize_t)ny, (hsize_t)nz};
    hid_t space = H5Screate_simple(3, dims, nullptr);
    if (space < 0) return false;
    hid_t ds = H5Dcreate2(file, name, type, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    bool ok = ds >= 0 && H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) >= 0;
    if (ds >= 0) ok = H5Dclose(ds) >= 0 && ok;
    return H5Sclose(space) >= 0 && ok;
}

bool write_1d(hid_t file, const char *name, const std::vector<double> &v) {
    hsize_t n = v.size();
    hid_t space = H5Screate_simple(1, &n, nullptr);
    if (space < 0) return false;
    hid_t ds = H5Dcreate2(file, name, H5T_NATIVE_DOUBLE, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    bool ok = ds >= 0 && H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data()) >= 0;
    if (ds >= 0) ok = H5Dclose(ds) >= 0 && ok;
    return H5Sclose(space) >= 0 && ok;
}

bool write_scalar(hid_t group, const char *name, hid_t type, const void *value) {
    hid_t space = H5Screate(H5S_SCALAR);
    if (space < 0) return false;
    hid_t ds = H5Dcreate2(group, name, type, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    bool ok = ds >= 0 && H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, value) >= 0;
    if (ds >= 0) ok = H5Dclose(ds) >= 0 && ok;
    return H5Sclose(space) >= 0 && ok;
}

bool failed(const char *what) {
    std::fprintf(stderr, "HDF5 error while %s\n", what);
    return false;
}

}  // namespace

bool write_h5(const char *path, const H5Result &r) {
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    hid_t file = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (file < 0) return failed("creating the output file");
    bool ok = write_field(file, "/mean", H5T_NATIVE_DOUBLE, r.nx, r.ny, r.nz, r.mean) &&
              write_field(file, "/variance", H5T_NATIVE_DOUBLE, r.nx, r.ny, r.nz, r.variance) &&
              write_field(file, "/standard_error", H5T_NATIVE_DOUBLE, r.nx, r.ny, r.nz,
                          r.standard_error) &&
              write_field(file, "/mean_steps", H5T_NATIVE_DOUBLE, r.nx, r.ny, r.nz, r.mean_steps) &&
              write_field(file, "/location", H5T_NATIVE_UCHAR, r.nx, r.ny, r.nz, r.location);
    if (!ok) {
        H5Fclose(file);
        return failed("writing field datasets");
    }
    std::vector<double> x(r.nx), y(r.ny), z(r.nz);
    for (int i = 0; i < r.nx; ++i) x[i] = grid_coordinate(r.bounds[0], r.bounds[1], i, r.nx);
    for (int i = 0; i < r.ny; ++i) y[i] = grid_coordinate(r.bounds[2], r.bounds[3], i, r.ny);
    for (int i = 0; i < r.nz; ++i) z[i] = grid_coordinate(r.bounds[4], r.bounds[5], i, r.nz);
    if (!write_1d(file, "/x", x) || !write_1d(file, "/y", y) || !write_1d(file, "/z", z)) {
        H5Fclose(file);
        return failed("writing coordinate datasets");
    }
    hid_t md = H5Gcreate2(file, "/metadata", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (md < 0) {
        H5Fclose(file);
        return failed("creating the metadata group");
    }
    ok = write_scalar(md, "N_walks", H5T_NATIVE_INT, &r.n_walks) &&
         write_scalar(md, "epsilon", H5T_NATIVE_DOUBLE, &r.epsilon) &&
         write_scalar(md, "max_steps", H5T_NATIVE_INT, &r.max_steps) &&
         write_scalar(md, "max_steps_hits", H5T_NATIVE_LLONG, &r.max_steps_hits) &&
         write_scalar(md, "seed", H5T_NATIVE_UINT64, &r.seed);
    if (ok && r.ray_metadata)
        ok = write_scalar(md, "max_ray_attempts", H5T_NATIVE_INT, &r.max_ray_attempts) &&
             write_scalar(md, "ambiguous_ray_retries", H5T_NATIVE_LLONG, &r.ambiguous_ray_retries) &&
             write_scalar(md, "indeterminate_points", H5T_NATIVE_LLONG, &r.indeterminate_points);
    if (ok && r.source_metadata)
        ok = write_scalar(md, "alpha", H5T_NATIVE_DOUBLE, &r.alpha) &&
             write_scalar(md, "source_mode", H5T_NATIVE_INT, &r.source_mode);
    const bool closed = H5Gclose(md) >= 0;
    if (!ok || !closed) {
        H5Fclose(file);
        return failed("writing metadata");
    }
    if (H5Fclose(file) < 0) return failed("closing the output file");
    return true;
}

}  // namespace woscli
