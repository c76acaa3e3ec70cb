/* A C consumer of include/picgpu.h, as a PETSc-side shim would be: prints the
 * size and field offsets of the ABI structs, then loads libpicgpu.so with
 * dlopen and calls two entry points that need no GPU (default options, and a
 * PETSc binary Vec file write).  Built and run by tests/test_abi.py.
 *   usage: abi_consumer <libpicgpu.so> <vec file to write> */
#include <dlfcn.h>
#include <stddef.h>
#include <stdio.h>

#include "picgpu.h"

#define FIELD(T, f) printf("%s.%s %zu %zu\n", #T, #f, offsetof(T, f), sizeof(((T *)0)->f))

int main(int argc, char **argv) {
  if (argc != 3) return 2;
  printf("sizeof picgpu_options %zu\n", sizeof(picgpu_options));
  FIELD(picgpu_options, precision);
  FIELD(picgpu_options, use_mlp_guess);
  FIELD(picgpu_options, max_traversal);
  FIELD(picgpu_options, distance_threshold);
  FIELD(picgpu_options, weight_clamp);
  FIELD(picgpu_options, bbox_tolerance);
  FIELD(picgpu_options, device);
  FIELD(picgpu_options, collect_walk_stats);
  printf("sizeof picgpu_stats %zu\n", sizeof(picgpu_stats));
  FIELD(picgpu_stats, n_particles);
  FIELD(picgpu_stats, located);
  FIELD(picgpu_stats, lost);
  FIELD(picgpu_stats, walk_iterations);
  FIELD(picgpu_stats, walk_max);
  FIELD(picgpu_stats, fp64_fallbacks);
  FIELD(picgpu_stats, removed_lost);
  FIELD(picgpu_stats, removed_out_of_bounds);
  FIELD(picgpu_stats, particle_steps);
  FIELD(picgpu_stats, walk_hist);
  printf("sizeof picgpu_boundary_config %zu\n", sizeof(picgpu_boundary_config));
  FIELD(picgpu_boundary_config, type);
  FIELD(picgpu_boundary_config, handler);
  FIELD(picgpu_boundary_config, velocity);
  FIELD(picgpu_boundary_config, vmax);

  void *lib = dlopen(argv[1], RTLD_NOW | RTLD_LOCAL);
  if (!lib) {
    fprintf(stderr, "dlopen: %s\n", dlerror());
    return 3;
  }
  void (*default_options)(picgpu_options *) = (void (*)(picgpu_options *))dlsym(lib, "picgpu_default_options");
  int (*write_vec)(const char *, int64_t, int, const double *) =
      (int (*)(const char *, int64_t, int, const double *))dlsym(lib, "picgpu_write_petsc_vec");
  if (!default_options || !write_vec) return 4;
  picgpu_options o;
  default_options(&o);
  printf("default precision %d use_mlp_guess %d max_traversal %d distance_threshold %.17g weight_clamp %.17g "
         "bbox_tolerance %.17g device %d collect_walk_stats %d\n",
         o.precision, o.use_mlp_guess, o.max_traversal, o.distance_threshold, o.weight_clamp, o.bbox_tolerance,
         o.device, o.collect_walk_stats);
  const double v[6] = {1.0, -2.5, 3.25, 0.0, 1e300, -1e-300};
  printf("write_petsc_vec %d\n", write_vec(argv[2], 6, 3, v));
  return 0;
}
