/* wos_run.h -- C entry point of the equation registry (libwos_run.so).
 *
 * The reference dispatches its CLI through a registry of
 *   struct Equation { const char *name; int (*run_2D)(...); int (*run_3D)(...); }
 * (src/equations/equations.hpp:7-31; lookup by exact name, src/main.cpp:285-304;
 * dispatch, src/main.cpp:380-389).  wos_hip_run is that call behind one C
 * symbol: the registry entry by index, the reference's run arguments
 * (rank 0 of 1: one process per GPU), an optional problem descriptor that
 * replaces the equation's built-in data, and the GPU execution options.  It
 * loads the OBJ mesh, solves on the grid over the mesh's bounding box, prints
 * the reference's progress and summary lines and writes the reference's
 * HDF5 layout (include/wos/hdf5_io.hpp:147-322).  Returns 0 on success and 1
 * on error, as the reference's run functions do.
 */
#ifndef WOS_RUN_H
#define WOS_RUN_H

#include <stdint.h>

#include "wos_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* GPU execution options of a registry run (NULL: FP64 plain, device 0, fixed
 * walks). */
typedef struct wos_exec_opts {
    int32_t precision;           /* WOS_FP64_PLAIN or WOS_FP32_PERSISTENT          */
    int32_t device;              /* HIP device index                                */
    double target_se;            /* > 0: adaptive walks per point until SE <= it    */
    int32_t max_walks;           /* adaptive cap per point (0: N_walks)             */
    int32_t reserved;
} wos_exec_opts;

/* Registry size and the name of entry `equation_id` ("laplace", "poisson",
 * "screened_poisson", "helmholtz"), NULL when out of range
 * (src/equations/equations.hpp:26-31). */
int wos_hip_equation_count(void);
const char *wos_hip_equation_name(int equation_id);

/* Equation::run_2D / run_3D (src/equations/equations.hpp:7-17) of registry
 * entry `equation_id` with dim = 2 or 3.  source_mode: 0 uniform, 1 green
 * (include/wos/source_mode.hpp:5-9).  problem: NULL for the reference's
 * built-in equation data.
 *
 * Threading: reentrant.  Every call creates and destroys its own GPU context
 * and keeps its execution options per thread, so calls may run concurrently
 * on several threads (on one GPU or several); their stdout lines interleave.
 * Two calls must not write the same output file.  The reference calls run
 * once per MPI rank, single-threaded. */
int wos_hip_run(int dim, const char *mesh_filename, const char *output_filename, int Nx, int Ny,
                int Nz, int N_walks, double epsilon, int max_steps, int max_ray_attempts,
                uint64_t seed, double alpha, int source_mode, int equation_id,
                const wos_eq_desc *problem, const wos_exec_opts *opts);

#ifdef __cplusplus
}
#endif

#endif /* WOS_RUN_H */
